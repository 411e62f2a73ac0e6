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
    # node valences 3-8: the gathers of an irregular triangulation
    "tri1m": ("cylinder_delaunay", 2, 0.15, 0.0, 6e6,
              "SST k-omega RANS FP64, cylinder, irregular Delaunay triangulation ~1M cells"),
    # 64 NACA 4-digit cases (seed 4004), per-case Mach and alpha, one batched context
    "batch64": ("batch", 2, None, None, 6e6, "SST k-omega RANS FP64, 64-case NACA 4-digit batch (seed 4004)"),
    # cylinder O-grid over a Cartesian background: overset coupling, dual time (BDF2, dt 0.2)
    "c2": ("overset", 1, 0.2, 0.0, 100.0, "laminar NS FP64, cylinder O-grid 256x96 over background 400x200, "
                                          "overset coupling, dual time"),
}


def build_mesh(kind):
    from tmp_amd import mesh
    if kind == "airfoil_1m":
        m = mesh.airfoil_1m()
    elif kind == "bump":
        m = mesh.channel_bump(seed=1001)
    elif kind == "cylinder_delaunay":
        m = mesh.cylinder_delaunay()
    else:
        raise ValueError(kind)
    return mesh.prepare(m)


def _batch_case_mesh(case):
    from tmp_amd import mesh
    return mesh.prepare(mesh.batch_mesh(case))[0]


def build_batch(workers=8):
    """The 64 batch cases and their prepared meshes (built in a process pool)."""
    from concurrent.futures import ProcessPoolExecutor
    from tmp_amd import mesh
    cases = mesh.batch_cases()
    with ProcessPoolExecutor(max_workers=workers) as ex:
        meshes = list(ex.map(_batch_case_mesh, cases))
    return cases, meshes


def host_cpu():
    """nproc and the CPU model name of this host (from /proc/cpuinfo)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def time_to_tolerance(n_train, n_test, epochs):
    """The metric's second half: wall time to ||R_rho/V|| = 1e-8 of the
    first residual on held-out laminar NACA 4-digit cases, for the plain
    cold start (each case alone, from the free stream), the batched cold
    start, and the batched NN warm start (inference and the cold-start
    screen included), with the network trained in this run on fields this
    solver converged (tmp_amd/warmstart.py)."""
    from tmp_amd import warmstart
    r = warmstart.experiment(n_train=n_train, n_test=n_test, epochs=epochs)
    t = r["measured_seconds_to_1e-8"]
    lv = r["levels"]["1e-08"]
    return {
        "family": r["family"], "train_cases": r["train_cases"], "test_cases": r["test_cases"],
        "epochs": r["epochs"], "train_seconds": r["train_seconds"], "tolerance": "1e-8 x cold first residual",
        "seconds_cold_plain": t["cold_plain_single_case_solves"], "seconds_cold_batched": t["cold_batched"],
        "seconds_warm_batched": t["warm_batched_incl_inference_and_screen"],
        "speedup_warm_batched_vs_cold_plain": r["measured_speedup_warm_batched_vs_cold_plain"],
        "speedup_warm_batched_vs_cold_batched": r["measured_speedup_warm_batched_vs_cold_batched"],
        "median_iteration_speedup": lv["median_iteration_speedup"],
        "warm_kept": r["warm_kept"], "fell_back_to_cold": r["fell_back_to_cold"],
        "max_abs_dCL": r["max_abs_dCL_converged"], "max_abs_dCD": r["max_abs_dCD_converged"],
    }


def mixed_error(solver, m, flow, mach, aoa, re):
    """One residual of the mixed-precision solver's final state in both
    precisions: max |R_mixed - R_fp64| / max |R_fp64| per equation (rho, rhoE,
    rhou, rhov, and rho k, rho omega for SST).  Mixed precision moves the
    near-cancelling k residual of SST states far more than the flow
    equations, so this is reported beside the mixed throughput."""
    import numpy as np
    from tmp_amd import usg
    ref = usg.Solver(m, flow=flow, mach=mach, aoa=aoa, reynolds=re)
    sst = flow == usg.SST
    ref.set_state(solver.U, solver.field(usg.F_CONS), solver.field(usg.F_TURB) if sst else None,
                  solver.field(usg.F_TCONS) if sst else None, solver.field(usg.F_MUT) if sst else None)
    solver.residual()
    ref.residual()
    a, b = solver.R, ref.R
    if sst:
        a = np.concatenate([a, solver.field(usg.F_TRES)], axis=1)
        b = np.concatenate([b, ref.field(usg.F_TRES)], axis=1)
    ref.close()
    return [float(x) for x in np.max(np.abs(a - b), axis=0) / np.maximum(np.max(np.abs(b), axis=0), 1e-300)]


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


def cpu_baseline_overset(seconds):
    """The oracle on the overset pair: its two serial solvers coupled by
    its stage-by-stage interpolation, dual time (BDF2, dt 0.2), sub-iterated
    until ``seconds`` of CPU work (1 core)."""
    from oracle import oracle as O
    from oracle.overset_oracle import OversetCoupling
    from tmp_amd import overset_c2
    og, bg, conn, _ = overset_c2.build()
    oms = [O.OracleMesh(og), O.OracleMesh(bg)]
    cp = OversetCoupling(oms, conn)
    ss = [O.OracleSolver(om, 1, overset_c2.MACH, 0.0, overset_c2.RE) for om in oms]
    for s, b in zip(ss, cp.blank):
        s.set_blank(b)
        s.dual_time(0.2, 2)
        s.dual_time_advance()
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        cp.apply(ss)
        for s in ss:
            s.lusgs_step()
        n += 1
        print(f"# cpu baseline sub-iteration {n}: {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    return (og.nc + bg.nc) * n / dt, n, dt


def _cpu_case(args):
    case, flow, re, seconds = args
    m = _batch_case_mesh(case)
    v, n, dt = cpu_baseline(m, flow, case["mach"], case["aoa"], re, seconds)
    return v, n, m.nc


def cpu_baseline_batch(cases, flow, re, seconds, workers):
    """One batch case per core (the oracle, 1 thread each) on ``workers``
    cores at once, each for ``seconds``: summed cell-iterations/s."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor
    sample = cases[:workers]
    # spawned (not forked) workers: this process already holds the GPU
    with ProcessPoolExecutor(max_workers=workers, mp_context=multiprocessing.get_context("spawn")) as ex:
        res = list(ex.map(_cpu_case, [(c, flow, re, seconds) for c in sample]))
    return sum(r[0] for r in res), sum(r[1] for r in res), len(sample)


def measured_traffic(kernel, workload, nc):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by what
    it reads for the plain copy kernel), or None when there is none for this
    mesh."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"{kernel}_{workload}_traffic.md")
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
    ap.add_argument("--no-time-to-tol", action="store_true",
                    help="skip the time-to-1e-8 comparison (NN warm vs plain cold) of the default workload")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="mixed: FP32 gradients/limiter/fluxes, FP64 state, accumulation, diagonal and norms")
    ap.add_argument("--autotune", action="store_true",
                    help="pick threads per block per kernel group and the face-cell launch scheme "
                         "(cached in ~/.cache/tmp_amd/autotune.json)")
    ap.add_argument("--face-cell", default="auto", choices=["auto", "separate", "fused"],
                    help="face-cell launch scheme (FP64): the library's default or the autotuner's choice (auto), "
                         "a face launch and a cell launch through HBM (separate), or one launch per tile with the "
                         "face record in LDS (fused)")
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
    mixed = a.precision == "mixed"
    prec = usg.MIXED if mixed else usg.FP64
    t = time.perf_counter()
    if kind == "overset":
        from tmp_amd import overset_c2
        og, bg, _, st = overset_c2.build()
        nc, nf = og.nc + bg.nc, og.nf + bg.nf
        print(f"# overset pair: {og.nc} + {bg.nc} cells, {len(st['recv'])} receptor and "
              f"{int((st['blank'] == 0).sum())} hole cells, built in {time.perf_counter() - t:.1f}s",
              file=sys.stderr, flush=True)
        solver = overset_c2.solver(og, bg, st, dt=0.2)
        solver.dual_time_advance()
        config = {"workload": label, "cells": nc, "faces": nf, "receptor_cells": int(len(st["recv"])),
                  "hole_cells": int((st["blank"] == 0).sum()), "mach": mach, "reynolds": re,
                  "dual_time": "BDF2, dt 0.2; steps are sub-iterations of one physical step"}
        metric = "cell-iters/s, overset cylinder (two grids), dual-time LU-SGS"
    elif kind == "batch":
        cases, meshes = build_batch()
        nc, nf = sum(m.nc for m in meshes), sum(m.nf for m in meshes)
        print(f"# batch of {len(meshes)} cases: {nc} cells, {nf} faces, built in {time.perf_counter() - t:.1f}s",
              file=sys.stderr, flush=True)
        solver = usg.BatchSolver(meshes, flow=flow, mach=[c["mach"] for c in cases],
                                 aoa=[c["aoa"] for c in cases], reynolds=re, precision=prec)
        config = {"workload": label, "cases": len(meshes), "cells": nc, "faces": nf,
                  "case_cells": [min(m.nc for m in meshes), max(m.nc for m in meshes)],
                  "mach": [0.3, 0.8], "aoa_deg": [-4.0, 8.0], "reynolds": re}
        metric = "cell-iters/s, 64-case SST batch"
    else:
        m, ncol = build_mesh(kind)
        nc, nf = m.nc, m.nf
        print(f"# mesh {m.name}: {nc} cells, {nf} faces, {ncol} colours, built in {time.perf_counter() - t:.1f}s",
              file=sys.stderr, flush=True)
        solver = usg.Solver(m, flow=flow, mach=mach, aoa=aoa, reynolds=re, precision=prec)
        config = {"workload": label, "cells": nc, "faces": nf, "colours": ncol, "mach": mach, "aoa_deg": aoa,
                  "reynolds": re}
        metric = "cell-iters/s at 1M-cell SST"
    if mixed:
        config["workload"] = config["workload"].replace("FP64", "mixed FP32/FP64")
    nfw, nbw = solver.waves()
    config.update({"lusgs_waves": [nfw, nbw], "integrator": "LU-SGS CFL 400", "parallelism": "single GPU"})
    if a.autotune:
        from tmp_amd import autotune
        t = time.perf_counter()
        cfg, hit = autotune.tune(solver, prec, nf)
        print(f"# autotune {'cache hit' if hit else 'search'}: launch config {cfg} ({time.perf_counter() - t:.1f}s)",
              file=sys.stderr, flush=True)
    if a.face_cell != "auto":
        solver.fused(a.face_cell == "fused")
    config["threads_per_block"] = solver.blocks()
    fused = solver.is_fused()
    config["face_cell_launch"] = "fused" if fused else "separate"
    if a.warmup:
        solver.lusgs(a.warmup)
    solver.sync()
    # timed region 1: the throughput, iterations as captured-graph replays
    t0 = time.perf_counter()
    hist = solver.lusgs(a.steps, history=True)   # returns after the stream drained
    t1 = time.perf_counter()
    elapsed = t1 - t0
    # timed region 2: the same iterations launched eagerly with HIP events
    # around every launch of the dominant kernel on its stream
    solver.timing(True)
    t2 = time.perf_counter()
    solver.lusgs(a.steps)
    t3 = time.perf_counter()
    ktime_ms, klaunch = solver.kernel_time()
    solver.timing(False)
    config["monitors"] = {"res_l2_first": float(hist.reshape(a.steps, -1)[0, 0]),
                          "res_l2_last": float(hist.reshape(a.steps, -1)[-1, 0]),
                          "sum_dW_rho_sq": float(solver.dw_rho_sq()[0])}
    if mixed and kind not in ("batch", "overset"):
        config["mixed_vs_fp64_residual"] = mixed_error(solver, m, flow, mach, aoa, re)
        print(f"# mixed vs FP64 residual of the final state: {config['mixed_vs_fp64_residual']}", file=sys.stderr,
              flush=True)
    if kind == "batch":
        st = solver.status()
        config["failed_cases"] = int((st != 0).sum())
        print(f"# cases frozen by a failed check: {config['failed_cases']} of {len(st)}", file=sys.stderr,
              flush=True)
    value = nc * a.steps / elapsed
    r = nf / nc
    # the dominant kernel: the face launch (or the fused face-cell launch)
    kname = "k_facecell" if fused else "k_face"
    kbytes = (roofline.fused_kernel_bytes_per_cell(flow, r) if fused
              else roofline.face_kernel_bytes_per_cell(flow, r, mixed)) * nc
    kavg_s = ktime_ms / 1e3 / max(klaunch, 1)
    achieved = kbytes / kavg_s / 1e9
    it_frac = roofline.iteration_bytes_per_cell(flow, r) * value / 1e9 / roofline.HBM_PEAK_GBS
    h = hist.reshape(a.steps, -1)
    print(f"# residual history (case 0) {h[0, 0]:.4e} -> {h[-1, 0]:.4e}; flux kernel avg {kavg_s * 1e6:.1f} us "
          f"over {klaunch} launches; iteration model {it_frac:.3f} of HBM peak; eager "
          f"{(t3 - t2) / a.steps * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
    out = {
        "metric": metric,
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
        "config": config,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved,
                     "peak": roofline.HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / roofline.HBM_PEAK_GBS,
                     "traffic": None if mixed else measured_traffic(kname, a.workload, nc),
                     "iteration_frac": it_frac},
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        nproc, model = host_cpu()
        if kind == "overset":
            v, n, dt = cpu_baseline_overset(a.cpu_seconds)
            out["cpu_baseline"] = {"value": v, "unit": "cell-iterations/s", "cores": 1, "kind": "port",
                                   "nproc": nproc, "cpu_model": model,
                                   "sample": f"{n} dual-time LU-SGS sub-iterations of the same two grids with the "
                                             f"oracle's coupling ({dt:.1f}s, oracle C restatement, -O2, 1 thread)"}
        elif kind == "batch":
            # one case per core, at most 16: the CPU share of a one-GPU host
            workers = min(16, nproc)
            v, n, k = cpu_baseline_batch(cases, flow, re, a.cpu_seconds, workers)
            out["cpu_baseline"] = {"value": v, "unit": "cell-iterations/s", "cores": workers, "kind": "port",
                                   "nproc": nproc, "cpu_model": model,
                                   "sample": f"first {k} batch cases, one per core, {n} LU-SGS iterations in "
                                             f"total over {a.cpu_seconds:.0f}s each (oracle C restatement, -O2)"}
        else:
            v, n, dt = cpu_baseline(m, flow, mach, aoa, re, a.cpu_seconds)
            out["cpu_baseline"] = {"value": v, "unit": "cell-iterations/s", "cores": 1, "kind": "port",
                                   "nproc": nproc, "cpu_model": model,
                                   "sample": f"{n} LU-SGS iterations of the same {nc}-cell workload from the "
                                             f"free stream ({dt:.1f}s, oracle C restatement, -O2, 1 thread)"}
    if a.workload == "sst1m" and not mixed and not a.no_time_to_tol:
        solver.close()
        print("# time to 1e-8: NN warm start vs plain cold start", file=sys.stderr, flush=True)
        out["time_to_tolerance"] = time_to_tolerance(256, 8, 150)
    print(json.dumps(out), flush=True)

if __name__ == "__main__":
    main()
