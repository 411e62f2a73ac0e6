"""Throughput of the gfx950 UnstructSolver path: LU-SGS cell-iterations/s.

Default workload: SST k-omega RANS, FP64, ~1M-cell hybrid mesh around NACA
0012 (quad wall layers with first spacing 4.5e-6 c, isotropic triangles to a
far field at 50 c), M = 0.15, alpha = 4 deg, Re = 6e6, cold start.  A step
is one LU-SGS pseudo-time iteration (residual + forward sweep + backward
sweep + update) over the whole mesh; the state is resident in HBM.

Prints one JSON line (see the keys below).  ``--gpus`` above 1 is refused:
this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # name: (mesh builder, flow, mach, aoa, reynolds, label)
    "sst1m": ("airfoil_1m", 2, 0.15, 4.0, 6e6, "SST k-omega RANS FP64, NACA0012 hybrid ~1M cells"),
    "euler1m": ("airfoil_1m", 0, 0.15, 4.0, 6e6, "Euler FP64, NACA0012 hybrid ~1M cells"),
    "bump20k": ("bump", 0, 0.5, 0.0, 1e6, "Euler FP64, channel bump ~20k cells"),
}


def build_mesh(kind):
    from tmp_amd import mesh
    if kind == "airfoil_1m":
        m = mesh.airfoil_1m()
    elif kind == "bump":
        m = mesh.channel_bump(seed=1001)
    else:
        raise ValueError(kind)
    return mesh.prepare(m)


def cpu_baseline(m, flow, mach, aoa, re, seconds):
    """The oracle (C restatement of the reference's serial path, 1 core) on
    the same mesh and inputs, iterated until ``seconds`` of CPU work."""
    from oracle import oracle as O
    print("# cpu baseline: building the oracle's geometry", file=sys.stderr, flush=True)
    om = O.OracleMesh(m)
    s = O.OracleSolver(om, flow, mach, aoa, re)
    n = 0
    t0 = time.perf_counter()
    while True:
        s.lusgs_step()
        n += 1
        print(f"# cpu baseline iteration {n}: {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return m.nc * n / dt, n, dt


def measured_traffic(workload, nc):
    """HBM bytes per k_face launch from the committed rocprofv3 --pmc summary
    of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by what it reads
    for the plain copy kernel), or None when there is none for this mesh."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"k_face_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and line.split(":")[0].strip() in ("cells", "traffic_bytes_per_launch"):
            k, v = line.split(":", 1)
            vals[k.strip()] = float(v.strip().strip("`"))
    if int(vals.get("cells", -1)) != nc or "traffic_bytes_per_launch" not in vals:
        return None
    return vals["traffic_bytes_per_launch"]


def copy_calibration(mb):
    """Runs the library's plain copy kernel (usg_copy_kernel) three times over
    ``mb`` MB: FETCH_SIZE / WRITE_SIZE of this dispatch are ``mb`` MB each."""
    import ctypes
    import torch
    from tmp_amd import usg
    n = mb * 1024 * 1024 // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    L = usg.lib()
    for _ in range(3):
        rc = L.usg_copy_kernel(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()),
                               ctypes.c_long(n * 8), None)
        if rc:
            raise RuntimeError(L.usg_last_error().decode())
    torch.cuda.synchronize()
    assert bool((dst == 1).all())
    print(f"# copy kernel: 3 x {mb} MB", file=sys.stderr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="sst1m", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="mixed: FP32 gradients/limiter/fluxes, FP64 state, accumulation, diagonal and norms")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run the plain copy kernel over MB megabytes (traffic-counter calibration)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1:
        print(f"bench.py: --gpus {a.gpus} is not supported; this build runs the path on one GPU "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    from tmp_amd import roofline, usg

    kind, flow, mach, aoa, re, label = WORKLOADS[a.workload]
    t = time.perf_counter()
    m, ncol = build_mesh(kind)
    print(f"# mesh {m.name}: {m.nc} cells, {m.nf} faces, {ncol} colours, built in {time.perf_counter() - t:.1f}s",
          file=sys.stderr, flush=True)
    mixed = a.precision == "mixed"
    solver = usg.Solver(m, flow=flow, mach=mach, aoa=aoa, reynolds=re, precision=usg.MIXED if mixed else usg.FP64)
    nfw, nbw = solver.waves()
    if a.warmup:
        solver.lusgs(a.warmup)
    solver.sync()
    # timed region 1: the throughput, iterations as captured-graph replays
    t0 = time.perf_counter()
    hist = solver.lusgs(a.steps, history=True)   # returns after the stream drained
    t1 = time.perf_counter()
    elapsed = t1 - t0
    # timed region 2: the same iterations launched eagerly with HIP events
    # around every launch of the dominant kernel (k_face) on its stream
    solver.timing(True)
    t2 = time.perf_counter()
    solver.lusgs(a.steps)
    t3 = time.perf_counter()
    ktime_ms, klaunch = solver.kernel_time()
    solver.timing(False)
    value = m.nc * a.steps / elapsed
    r = m.nf / m.nc
    kbytes = roofline.face_kernel_bytes_per_cell(flow, r, mixed) * m.nc
    kavg_s = ktime_ms / 1e3 / max(klaunch, 1)
    achieved = kbytes / kavg_s / 1e9
    it_frac = roofline.iteration_bytes_per_cell(flow, r) * value / 1e9 / roofline.HBM_PEAK_GBS
    print(f"# residual history {hist[0]:.4e} -> {hist[-1]:.4e}; flux kernel avg {kavg_s * 1e6:.1f} us over "
          f"{klaunch} launches; iteration model {it_frac:.3f} of HBM peak; eager "
          f"{(t3 - t2) / a.steps * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
    out = {
        "metric": "cell-iters/s at 1M-cell SST",
        "value": value,
        "unit": "cell-iterations/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic: generated mesh, free-stream cold start",
        "config": {"workload": label.replace("FP64", "mixed FP32/FP64") if mixed else label, "cells": m.nc, "faces": m.nf, "colours": ncol,
                   "lusgs_waves": [nfw, nbw], "mach": mach, "aoa_deg": aoa, "reynolds": re,
                   "integrator": "LU-SGS CFL 400", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_face", "achieved": achieved,
                     "peak": roofline.HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / roofline.HBM_PEAK_GBS,
                     "traffic": None if mixed else measured_traffic(a.workload, m.nc),
                     "iteration_frac": it_frac},
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        v, n, dt = cpu_baseline(m, flow, mach, aoa, re, a.cpu_seconds)
        out["cpu_baseline"] = {"value": v, "unit": "cell-iterations/s", "cores": 1, "kind": "port",
                               "sample": f"{n} LU-SGS iterations of the same {m.nc}-cell workload from the "
                                         f"free stream ({dt:.1f}s, oracle C restatement, -O2, 1 thread)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
