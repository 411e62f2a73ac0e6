"""Learned warm start for the pseudo-time solver.

A coordinate MLP (8 linear layers x 256, SiLU; include/usg_mlp.h, BF16 MFMA
on gfx950) maps a cell's position and the case's conditioning to its
converged primitive state; the solver then starts from that field
(usg_set_prim) instead of the free stream (Euler2D::initializeData,
Euler2D.cpp:153-171) and iterates with the unchanged LU-SGS path.

Family: laminar flow around NACA 4-digit airfoils (thickness t, camber m at
p = 0.4), Mach M, angle of attack alpha, Reynolds number Re, each case on its
own hybrid O-mesh (96 contour points, first spacing 1e-3 c, far field 20 c).

Inputs per cell (73): x^ = 2 (x - 0.5) / (1 + r), y^ = 2 y / (1 + r) with r
the distance to mid-chord; sin and cos of 16 frequencies 2^(j/2) pi/2 of x^
and of y^; log10 of the wall distance; the conditioning (M, alpha, log Re,
t, m, p).  Outputs: the state relative to the free stream, scaled to O(1):
(rho - 1) / M^2, u / (M sqrt(gamma)) - cos(alpha), v / (M sqrt(gamma)) -
sin(alpha), (p - 1) / (gamma M^2).

Run ``python -m tmp_amd.warmstart`` for the cold-versus-warm experiment
(prints progress and one JSON summary line); ``--checkpoints DIR`` keeps
each batch of training solves as a checkpoint and reuses it on the next run.
``--save-net FILE`` writes the trained network (weights, biases and the
training figures, .npz) and ``--load-net FILE`` starts from such a file
instead of building the dataset and training, so the two halves of a large
run can go in separate processes.
"""
from __future__ import annotations

import argparse
import json
import math
import multiprocessing
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from . import mesh, usg
from . import mlp as _mlp

GAMMA = 1.4
NFREQ = 16
NIN = 4 * NFREQ + 3 + 6
NOUT = 4
FAMILY_MESH = dict(ni=96, dy1=1e-3, rfar=20.0)


def family_cases(n: int, seed: int) -> list[dict]:
    rng = np.random.default_rng(seed)
    return [{"t": float(rng.uniform(0.08, 0.16)), "m": float(rng.uniform(0.0, 0.03)), "p": 0.4,
             "mach": float(rng.uniform(0.2, 0.5)), "aoa": float(rng.uniform(-2.0, 6.0)),
             "re": float(10 ** rng.uniform(3.0, 3.5))} for _ in range(n)]


def family_mesh(case: dict) -> mesh.Mesh:
    return mesh.prepare(mesh.airfoil_hybrid(t=case["t"], camber=case["m"], **FAMILY_MESH))[0]


def meshes_for(cases, workers=8):
    # spawned (not forked) workers: this process may already hold the GPU
    with ProcessPoolExecutor(max_workers=workers, mp_context=multiprocessing.get_context("spawn")) as ex:
        return list(ex.map(family_mesh, cases))


def features(m: mesh.Mesh, case: dict) -> np.ndarray:
    g = usg.mesh_geometry(m)
    cc, wd = g["cc"], g["wdist"]
    xr, y = cc[:, 0] - 0.5, cc[:, 1]
    r = np.hypot(xr, y)
    xh, yh = 2 * xr / (1 + r), 2 * y / (1 + r)
    cols = [xh, yh, np.log10(wd + 1e-5) / 5 + 1]
    for j in range(NFREQ):
        f = 2 ** (j / 2) * math.pi / 2
        cols += [np.sin(f * xh), np.cos(f * xh), np.sin(f * yh), np.cos(f * yh)]
    cond = [case["mach"] * 2, case["aoa"] / 10, math.log10(case["re"]) - 3, case["t"] * 10, case["m"] * 50,
            case["p"]]
    cols += [np.full(m.nc, v) for v in cond]
    return np.stack(cols, axis=1).astype(np.float32)


def targets(U: np.ndarray, case: dict) -> np.ndarray:
    """Primitive U[nc] = (rho, p, u, v) -> network outputs."""
    M, a = case["mach"], math.radians(case["aoa"])
    q = M * math.sqrt(GAMMA)
    return np.stack([(U[:, 0] - 1) / M ** 2, U[:, 2] / q - math.cos(a), U[:, 3] / q - math.sin(a),
                     (U[:, 1] - 1) / (GAMMA * M ** 2)], axis=1).astype(np.float32)


def fields(T: np.ndarray, case: dict) -> np.ndarray:
    """Network outputs -> primitive U (density and pressure kept positive)."""
    M, a = case["mach"], math.radians(case["aoa"])
    q = M * math.sqrt(GAMMA)
    T = np.asarray(T, np.float64)
    rho = np.maximum(1 + M ** 2 * T[:, 0], 0.05)
    p = np.maximum(1 + GAMMA * M ** 2 * T[:, 3], 0.05)
    return np.stack([rho, p, q * (T[:, 1] + math.cos(a)), q * (T[:, 2] + math.sin(a))], axis=1)


def batch_solver(meshes, cases) -> usg.BatchSolver:
    return usg.BatchSolver(meshes, flow=usg.LAMINAR, mach=[c["mach"] for c in cases],
                           aoa=[c["aoa"] for c in cases], reynolds=[c["re"] for c in cases])


def iterate_to(solver, tol: np.ndarray, max_iters: int, chunk: int = 500, log=None):
    """LU-SGS until every case's ||R_rho/V|| is <= tol[case] (or max_iters):
    per case, the first iteration (1-based) that met it (0 = never), and the
    full history [iters, ncase]."""
    hist = []
    first = np.zeros(solver.ncase, np.int64)
    done = 0
    while done < max_iters:
        h = solver.lusgs(min(chunk, max_iters - done), history=True)
        for i in range(solver.ncase):
            if first[i] == 0:
                hit = np.nonzero(h[:, i] <= tol[i])[0]
                if hit.size:
                    first[i] = done + hit[0] + 1
        hist.append(h)
        done += h.shape[0]
        if log:
            log(f"  {done} iterations, {int((first > 0).sum())}/{solver.ncase} cases at tolerance")
        if np.all(first > 0):
            break
    return first, np.concatenate(hist, axis=0)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _same_cases(s, cs) -> bool:
    return len(s.case_params) == len(cs) and all(
        p["mach"] == c["mach"] and p["aoa"] == c["aoa"] and p["reynolds"] == c["re"]
        for p, c in zip(s.case_params, cs))


def make_dataset(cases, iters, keep_drop=1e-5, batch=64, checkpoint_dir=None):
    """Converged fields of the cases (laminar, batched solves of `iters`
    iterations); cases that failed or whose residual did not drop by
    keep_drop are left out.  With checkpoint_dir, each solved batch is saved
    there (Solver.save) and a later call loads it instead of solving again."""
    X, Y, kept = [], [], []
    for a in range(0, len(cases), batch):
        cs = cases[a:a + batch]
        t0 = time.perf_counter()
        path = os.path.join(checkpoint_dir, f"train_{a:05d}_{len(cs)}.npz") if checkpoint_dir else None
        if path and os.path.exists(path):
            s = usg.load_checkpoint(path)
            if not _same_cases(s, cs) or s.step != iters:
                raise ValueError(f"{path} holds other cases or iterations; remove it to solve again")
            ms = s.meshes
        else:
            ms = meshes_for(cs)
            s = batch_solver(ms, cs)
            iterate_to(s, np.full(len(cs), -1.0), iters, chunk=1000)
            if path:
                os.makedirs(checkpoint_dir, exist_ok=True)
                s.save(path)
        h = s.hist
        st = s.status()
        U = s.U
        for i, (m, c) in enumerate(zip(ms, cs)):
            drop = h[-1, i] / h[0, i]
            if st[i] == 0 and drop <= keep_drop:
                X.append(features(m, c))
                Y.append(targets(U[s.case_slice(i)], c))
                kept.append(a + i)
        log(f"dataset: cases {a}..{a + len(cs) - 1} solved in {time.perf_counter() - t0:.1f}s, kept {len(kept)}")
        s.close()
    return np.concatenate(X), np.concatenate(Y), kept


def _init(nin, width, nlayer, nout, seed):
    """Uniform +-sqrt(6/fan_in) weights (+-sqrt(3/fan_in) on the linear output
    layer), zero biases; layout of include/usg_mlp.h."""
    rng = np.random.default_rng(seed)
    w, b = [], []
    for l in range(nlayer):
        k = nin if l == 0 else width
        n = nout if l == nlayer - 1 else width
        lim = math.sqrt((3.0 if l == nlayer - 1 else 6.0) / k)
        w.append(rng.uniform(-lim, lim, size=k * n))
        b.append(np.zeros(n))
    return np.concatenate(w).astype(np.float32), np.concatenate(b).astype(np.float32)


def train_net(X, Y, epochs, batch=16384, lr0=2e-3, lr1=2e-5, seed_init=5006, seed_perm=5005):
    net = _mlp.MLP(NIN, 256, 8, NOUT)
    net.set_params(*_init(NIN, 256, 8, NOUT, seed_init))
    net.set_data(X, Y)
    rng = np.random.default_rng(seed_perm)
    losses = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        lr = lr1 + 0.5 * (lr0 - lr1) * (1 + math.cos(math.pi * ep / max(1, epochs - 1)))
        losses.append(net.train_epoch(rng.permutation(len(X)), batch, lr))
        log(f"epoch {ep}: loss {losses[-1]:.4e} lr {lr:.2e} ({time.perf_counter() - t0:.1f}s)")
    return net, losses, time.perf_counter() - t0


def screen_prediction(U: np.ndarray) -> bool:
    """A predicted primitive field can start a solve only if it is finite
    with positive density and pressure everywhere."""
    U = np.asarray(U)
    return bool(np.all(np.isfinite(U)) and np.all(U[:, 0] > 0) and np.all(U[:, 1] > 0))


def screened_start(solver, U_pred: np.ndarray, probe: int = 20, growth: float = 10.0):
    """Warm start with the plain cold start as fallback and screen, per case.

    The cold start (the free stream, Euler2D.cpp:153-171) fixes the bar: its
    first residual r_cold.  A case starts from its prediction if the
    prediction is finite with positive density and pressure, and is screened
    after ``probe`` LU-SGS iterations: it keeps the warm start if every
    residual of the probe is finite, the case passed the solver's checks,
    and its last residual is below both r_cold and ``growth`` times its own
    first one.  Every other case restarts from the free stream.  Returns
    (warm[ncase] booleans, r_cold[ncase], probe iterations taken)."""
    nc = solver.ncase
    cell = solver.case_cell
    solver.init_freestream()
    solver.residual()
    r_cold = np.atleast_1d(solver.res_l2()).astype(np.float64)
    fs = np.concatenate([np.tile(solver.freestream(i), (int(cell[i + 1] - cell[i]), 1)) for i in range(nc)])
    U_pred = np.asarray(U_pred, np.float64)
    warm = np.array([screen_prediction(U_pred[cell[i]:cell[i + 1]]) for i in range(nc)])
    U0 = fs.copy()
    for i in np.nonzero(warm)[0]:
        U0[cell[i]:cell[i + 1]] = U_pred[cell[i]:cell[i + 1]]
    solver.set_prim(U0)
    if not warm.any() or probe <= 0:
        return warm, r_cold, 0
    try:
        h = np.asarray(solver.lusgs(probe, history=True)).reshape(probe, nc)
        st = solver.status() if isinstance(solver, usg.BatchSolver) else np.zeros(1, np.int32)
    except usg.UsgError:
        # a single-case context fails the call when a check breaks
        h = np.full((probe, nc), np.nan)
        st = np.ones(nc, np.int32)
    ok = np.all(np.isfinite(h), axis=0) & (st == 0) & (h[-1] < r_cold) & (h[-1] < growth * h[0])
    keep = warm & ok
    if np.any(warm & ~keep):
        U = solver.U
        for i in np.nonzero(~keep)[0]:
            U[cell[i]:cell[i + 1]] = fs[cell[i]:cell[i + 1]]
        solver.set_prim(U)
    return keep, r_cold, probe


def first_hits(hist: np.ndarray, tol: np.ndarray) -> np.ndarray:
    """Per case, the first iteration (1-based) whose norm is <= tol (0 = never)."""
    out = np.zeros(hist.shape[1], np.int64)
    for i in range(hist.shape[1]):
        hit = np.nonzero(hist[:, i] <= tol[i])[0]
        out[i] = hit[0] + 1 if hit.size else 0
    return out


LEVELS = (1e-4, 1e-6, 1e-8)


def save_net(path, net, n_train, n_kept, n_points, epochs, losses, t_train):
    """The trained network and its training figures as one .npz."""
    w, b = net.get_params()
    np.savez(path, w=w, b=b, losses=np.asarray(losses, np.float64),
             meta=np.array([n_train, n_kept, n_points, epochs], np.int64), t_train=np.float64(t_train))


def load_net(path):
    """(net, losses, t_train, (n_train, n_kept, n_points, epochs)) from save_net's file."""
    with np.load(path, allow_pickle=False) as z:
        net = _mlp.MLP(NIN, 256, 8, NOUT)
        net.set_params(z["w"], z["b"])
        return net, z["losses"].tolist(), float(z["t_train"]), tuple(int(v) for v in z["meta"])


def experiment(n_train=256, n_test=16, train_iters=20000, epochs=150, max_iters=40000, keep_drop=1e-5,
               checkpoint_dir=None, save_path=None, load_path=None, train_only=False):
    t_all = time.perf_counter()
    test_cases = family_cases(n_test, 5007)
    if load_path:
        net, losses, t_train, (n_train, n_kept, n_points, epochs) = load_net(load_path)
        log(f"network from {load_path}: {n_kept} of {n_train} training fields, {epochs} epochs")
    else:
        train_cases = family_cases(n_train, 5005)
        X, Y, kept = make_dataset(train_cases, train_iters, keep_drop, checkpoint_dir=checkpoint_dir)
        log(f"dataset: {len(kept)} fields, {len(X)} points")
        net, losses, t_train = train_net(X, Y, epochs)
        n_kept, n_points = len(kept), int(len(X))
        del X, Y
        if save_path:
            save_net(save_path, net, n_train, n_kept, n_points, epochs, losses, t_train)
            log(f"network saved to {save_path}")
    if train_only:
        return {"train_cases": n_kept, "train_points": n_points, "epochs": epochs, "train_seconds": t_train,
                "final_loss": losses[-1], "saved_to": save_path, "total_seconds": time.perf_counter() - t_all}
    ms = meshes_for(test_cases)
    # cold: free stream; the tolerances are relative to its first residual
    cold = batch_solver(ms, test_cases)
    h0 = cold.lusgs(1, history=True)[0]
    cold.init_freestream()
    tol = LEVELS[-1] * h0
    cold.sync()
    t0 = time.perf_counter()
    n_cold, h_cold = iterate_to(cold, tol, max_iters, log=log)
    t_cold_batch = time.perf_counter() - t0
    cl_c, cd_c = cold.forces()
    # plain: each case alone in its own context from the free stream
    t_plain = plain_solves(ms, test_cases, tol, max_iters)
    # warm: network prediction, screened against the cold start
    warm = batch_solver(ms, test_cases)
    warm.sync()
    t0 = time.perf_counter()
    U0 = np.concatenate([fields(net.predict(features(m, c)), c) for m, c in zip(ms, test_cases)])
    t_pred = time.perf_counter() - t0
    warm_kept, _, nprobe = screened_start(warm, U0)
    n_warm, h_warm = iterate_to(warm, tol, max_iters, log=log)
    t_warm_batch = time.perf_counter() - t0
    n_warm = np.where(n_warm > 0, n_warm + nprobe, 0)
    h_warm = np.concatenate([warm.hist[:nprobe], h_warm]) if nprobe else h_warm
    cl_w, cd_w = warm.forces()
    # single-case wall time per iteration (the time-to-tolerance of one solve)
    s = usg.Solver(ms[0], flow=usg.LAMINAR, mach=test_cases[0]["mach"], aoa=test_cases[0]["aoa"],
                   reynolds=test_cases[0]["re"])
    s.lusgs(50)
    s.sync()
    t0 = time.perf_counter()
    s.lusgs(1000)
    ms_it = (time.perf_counter() - t0) / 1000 * 1e3
    pred_ms = 1e3 * t_pred / n_test
    levels = {}
    for lv in LEVELS:
        nc_, nw_ = first_hits(h_cold, lv * h0), first_hits(h_warm, lv * h0)
        both = (nc_ > 0) & (nw_ > 0)
        sp = nc_[both] / nw_[both]
        tsp = (nc_[both] * ms_it) / (nw_[both] * ms_it + pred_ms)
        levels[f"{lv:g}"] = {"iters_cold": nc_.tolist(), "iters_warm": nw_.tolist(), "cases_both": int(both.sum()),
                             "median_iteration_speedup": float(np.median(sp)) if sp.size else None,
                             "median_time_speedup": float(np.median(tsp)) if tsp.size else None}
    both = (n_cold > 0) & (n_warm > 0)
    res = {
        "family": "laminar NACA 4-digit, M 0.2-0.5, alpha -2..6, Re 1e3-3.2e3, 96-point hybrid mesh",
        "train_cases": n_kept, "train_points": n_points, "epochs": epochs, "train_seconds": t_train,
        "final_loss": losses[-1], "test_cases": n_test, "max_iters": max_iters,
        "tolerances": "x the cold run's first residual", "levels": levels,
        "initial_residual_warm_over_cold": float(np.median(h_warm[0] / h_cold[0])),
        "ms_per_iteration_single_case": ms_it, "predict_ms_per_case": pred_ms,
        "max_abs_dCL_converged": float(np.max(np.abs(cl_w - cl_c)[both])) if both.any() else None,
        "max_abs_dCD_converged": float(np.max(np.abs(cd_w - cd_c)[both])) if both.any() else None,
        "warm_kept": int(warm_kept.sum()), "fell_back_to_cold": int((~warm_kept).sum()),
        "measured_seconds_to_1e-8": {
            "cold_plain_single_case_solves": t_plain,
            "cold_batched": t_cold_batch,
            "warm_batched_incl_inference_and_screen": t_warm_batch},
        "measured_speedup_warm_batched_vs_cold_batched": t_cold_batch / t_warm_batch,
        "measured_speedup_warm_batched_vs_cold_plain": t_plain / t_warm_batch,
        "total_seconds": time.perf_counter() - t_all,
    }
    return res


def plain_solves(ms, cases, tol, max_iters, chunk=500):
    """Wall time of solving each case alone in a single-case context from
    the free stream (the plain method), one after the other, each until its
    ||R_rho/V|| reaches tol[case] (or max_iters)."""
    total = 0.0
    for i, (m, c) in enumerate(zip(ms, cases)):
        s = usg.Solver(m, flow=usg.LAMINAR, mach=c["mach"], aoa=c["aoa"], reynolds=c["re"])
        s.sync()
        t0 = time.perf_counter()
        done = 0
        while done < max_iters:
            h = s.lusgs(min(chunk, max_iters - done), history=True)
            done += len(h)
            if np.any(h <= tol[i]):
                break
        total += time.perf_counter() - t0
        s.close()
        log(f"  plain solve {i + 1}/{len(cases)}: {done} iterations")
    return total


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=256)
    ap.add_argument("--test", type=int, default=16)
    ap.add_argument("--train-iters", type=int, default=20000)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--max-iters", type=int, default=40000)
    ap.add_argument("--checkpoints", default=None, help="directory of training-solve checkpoints (reused)")
    ap.add_argument("--save-net", default=None, help="write the trained network to this .npz")
    ap.add_argument("--load-net", default=None, help="start from this trained network (no dataset, no training)")
    ap.add_argument("--train-only", action="store_true", help="stop after training (with --save-net)")
    a = ap.parse_args()
    r = experiment(a.train, a.test, a.train_iters, a.epochs, a.max_iters, checkpoint_dir=a.checkpoints,
                   save_path=a.save_net, load_path=a.load_net, train_only=a.train_only)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
