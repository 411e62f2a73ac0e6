"""Summarise a rocprofv3 results database (``*_results.db``) as text: per
kernel call count, total and average duration, share of GPU time.

    python -m tmp_amd.profsummary <results.db> [--title T] [--command C]
    python -m tmp_amd.profsummary --pmc <results.db>
    python -m tmp_amd.profsummary --traffic KERNEL --cells NC --copy-bytes B \
        <fetch.db> <write.db> <copy_fetch.db> <copy_write.db>
    python -m tmp_amd.profsummary --sq <trace.db> <sq_pmc.db> --kernels K...

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
        if cn.startswith("GRBM_"):   # whole-GPU counters: the same value per instance, take the max
            acc[name][cn][d] = max(acc[name][cn].get(d, 0.0), v)
        else:                        # per-SE / per-channel counters: sum the instances
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


def hbm(trace_db: str, dbs: list, copy_bytes: int, kernels: list, command: str) -> None:
    """Per kernel: average duration (kernel trace), HBM bytes per dispatch
    (FETCH_SIZE and WRITE_SIZE, each scaled by its plain-copy calibration)
    and the achieved GB/s against the 8 TB/s peak."""
    fetch_db, write_db, cfetch_db, cwrite_db = dbs
    sf = copy_bytes / (_one(cfetch_db, "k_copy", "FETCH_SIZE") * 1024.0)
    sw = copy_bytes / (_one(cwrite_db, "k_copy", "WRITE_SIZE") * 1024.0)
    avg = {name: a for name, n, tot, a, pct in summary(trace_db)}
    fc, wc = counters(fetch_db), counters(write_db)
    print("# HBM bytes and achieved bandwidth per kernel\n")
    if command:
        print(f"Commands: {command}\n")
    print(f"FETCH_SIZE scale {sf:.4f}, WRITE_SIZE scale {sw:.4f} (plain copy of {copy_bytes} B).\n")
    print("| kernel | avg us | read MB | written MB | GB/s | of 8 TB/s |")
    print("|---|---:|---:|---:|---:|---:|")
    for k in kernels:
        names = [n for n in avg if k in n]
        if not names:
            continue
        n = names[0]
        f = [v["FETCH_SIZE"] for kk, v in fc.items() if k in kk and "FETCH_SIZE" in v]
        w = [v["WRITE_SIZE"] for kk, v in wc.items() if k in kk and "WRITE_SIZE" in v]
        if not f or not w:
            continue
        rb, wb = f[0] * 1024.0 * sf, w[0] * 1024.0 * sw
        gbs = (rb + wb) / (avg[n] * 1e-6) / 1e9
        print(f"| `{n[:70]}` | {avg[n]:.1f} | {rb / 1e6:.1f} | {wb / 1e6:.1f} | {gbs:.0f} | {gbs / 8000:.2f} |")


def sq(trace_db: str, pmc_db: str, kernels: list, command: str, simds: int = 1024, ghz: float = 2.4) -> None:
    """Per kernel, per wave: VALU instructions, vector memory reads and
    writes, the wave's lifetime and the share of it spent waiting to issue
    (SQ_WAVES, SQ_INSTS_VALU, SQ_INSTS_VMEM_RD/WR, SQ_WAVE_CYCLES,
    SQ_WAIT_INST_ANY); and the share of the chip's VALU issue cycles the
    kernel's VALU instructions take over its average duration (4 cycles per
    wave64 instruction on a 16-lane SIMD, `simds` SIMDs at `ghz`)."""
    times = {name: avg for name, _, _, avg, _ in summary(trace_db)}
    cs = counters(pmc_db)
    print(f"# SQ instruction and wait counters per kernel\n")
    if command:
        print(f"Commands: {command}\n")
    print("| kernel | waves | VALU / wave | loads / wave | stores / wave | wave cycles / wave | waiting | "
          f"VALU issue share ({simds} SIMDs, {ghz} GHz) |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|")
    for kn in kernels:
        hits = [(k, v) for k, v in cs.items() if kn in k]
        if len(hits) != 1:
            raise SystemExit(f"{pmc_db}: {len(hits)} kernels match {kn!r}")
        k, v = hits[0]
        w = v["SQ_WAVES"]
        tk = [t for name, t in times.items() if name == k]
        share = (w * v["SQ_INSTS_VALU"] / w * 4) / (simds * tk[0] * 1e-6 * ghz * 1e9) if tk else float("nan")
        print(f"| `{k[:60]}` | {w:.0f} | {v['SQ_INSTS_VALU'] / w:.0f} | {v['SQ_INSTS_VMEM_RD'] / w:.1f} | "
              f"{v['SQ_INSTS_VMEM_WR'] / w:.1f} | {v['SQ_WAVE_CYCLES'] / w:.0f} | "
              f"{v['SQ_WAIT_INST_ANY'] / v['SQ_WAVE_CYCLES']:.2f} | {share:.2f} |")


def mfma(trace_db: str, pmc_db: str, command: str, peak_tflops: float = 2500.0) -> None:
    """Per MFMA kernel: BF16 FLOPs per dispatch (SQ_INSTS_VALU_MFMA_MOPS_BF16 x 512),
    average duration, TFLOP/s and its fraction of the dense BF16 peak."""
    avg = {name: a for name, n, tot, a, pct in summary(trace_db)}
    cs = counters(pmc_db)
    print("# MFMA utilisation of the warm-start MLP kernels\n")
    if command:
        print(f"Commands: {command}\n")
    print("| kernel | avg us | GFLOP per dispatch | TFLOP/s | of 2.5 PFLOP/s |")
    print("|---|---:|---:|---:|---:|")
    for k, v in sorted(cs.items()):
        if "SQ_INSTS_VALU_MFMA_MOPS_BF16" not in v or v["SQ_INSTS_VALU_MFMA_MOPS_BF16"] <= 0:
            continue
        names = [n for n in avg if n.split("(")[0] == k.split("(")[0]]
        if not names:
            continue
        flop = v["SQ_INSTS_VALU_MFMA_MOPS_BF16"] * 512
        t = avg[names[0]] * 1e-6
        print(f"| `{k[:60]}` | {avg[names[0]]:.1f} | {flop / 1e9:.3f} | {flop / t / 1e12:.1f} | "
              f"{flop / t / 1e12 / peak_tflops:.3f} |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db", nargs="+")
    ap.add_argument("--hbm", metavar="TRACE_DB", help="achieved GB/s per kernel: db = fetch write copy_fetch copy_write")
    ap.add_argument("--kernels", nargs="*", default=["k_face", "k_cell", "k_grad", "k_sweep<2, true",
                                                     "k_sweep<2, false", "k_norm_partial"])
    ap.add_argument("--mfma", metavar="TRACE_DB", help="MFMA utilisation: db = the --pmc db")
    ap.add_argument("--sq", metavar="TRACE_DB", help="SQ counters per wave: db = the SQ --pmc db")
    ap.add_argument("--traffic", metavar="KERNEL")
    ap.add_argument("--cells", type=int, default=0)
    ap.add_argument("--copy-bytes", type=int, default=0)
    ap.add_argument("--title", default="rocprofv3 kernel summary")
    ap.add_argument("--command", default="")
    ap.add_argument("--pmc", action="store_true", help="the db holds --pmc counters")
    a = ap.parse_args()
    if a.hbm:
        hbm(a.hbm, a.db, a.copy_bytes, a.kernels, a.command)
        return
    if a.mfma:
        mfma(a.mfma, a.db[0], a.command)
        return
    if a.sq:
        sq(a.sq, a.db[0], a.kernels, a.command)
        return
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
