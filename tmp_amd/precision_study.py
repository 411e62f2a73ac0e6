"""Mixed precision against FP64 at the 1M-cell SST configuration.

For each angle of attack, an FP64 and a mixed-precision context (USG_MIXED:
FP32 gradients, limiter and face fluxes; FP64 state, residual sums, LU-SGS
diagonal and norms) on the same ~1M-cell hybrid NACA 0012 mesh (M 0.15,
Re 6e6, SST k-omega):

* trajectory: both run the same number of LU-SGS iterations from the free
  stream; per field (rho, u, v, p, k, omega) max |mixed - fp64| / max |fp64|
  and RMS(mixed - fp64) / RMS(fp64), and the C_L / C_D differences;
* same state: both evaluate one residual of the FP64 run's final state; per
  equation (rho, rho E, rho u, rho v, rho k, rho omega) the same two
  measures, and the C_L / C_D differences of that state.

The SST start-up diverges after ~230 iterations at the reference's CFL 400
on this mesh, so the default stays inside that window (150 iterations).
With ``--order 1`` (first-order reconstruction, usg_set_order) SST
converges on this mesh, and a long run compares the two precisions' nearly
converged solutions instead of a transient.

    python -m tmp_amd.precision_study [--iters N] [--aoa 0 2 4 ...] [--order 1|2]

Prints a progress line per case and one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import mesh, usg

FIELDS = ("rho", "u", "v", "p", "k", "omega")


def fields_of(s: usg.Solver) -> np.ndarray:
    U = s.U                      # (rho, p, u, v)
    T = s.field(usg.F_TURB)      # (k, omega)
    return np.stack([U[:, 0], U[:, 2], U[:, 3], U[:, 1], T[:, 0], T[:, 1]], axis=1)


def _rel(d, a):
    return ((np.max(np.abs(d), axis=0) / np.max(np.abs(a), axis=0)).tolist(),
            (np.sqrt(np.mean(d * d, axis=0)) / np.sqrt(np.mean(a * a, axis=0))).tolist())


EQS = ("rho", "rhoE", "rhou", "rhov", "rhok", "rhoomega")


def run(iters: int, aoas, mach=0.15, re=6e6, order=2):
    m, _ = mesh.prepare(mesh.airfoil_1m())
    rows = []
    for aoa in aoas:
        ctx = {}
        for prec in (usg.FP64, usg.MIXED):
            s = usg.Solver(m, flow=usg.SST, mach=mach, aoa=aoa, reynolds=re, precision=prec)
            s.set_order(order)
            s.sync()
            t0 = time.perf_counter()
            h = s.lusgs(iters, history=True)
            dt = time.perf_counter() - t0
            ctx[prec] = (s, fields_of(s), s.forces(), h, m.nc * iters / dt)
        a, b = ctx[usg.FP64][1], ctx[usg.MIXED][1]
        mx, rms = _rel(b - a, a)
        row = {"aoa": aoa,
               "trajectory": {"max_rel": dict(zip(FIELDS, mx)), "rms_rel": dict(zip(FIELDS, rms)),
                              "dCL": abs(ctx[usg.MIXED][2][0] - ctx[usg.FP64][2][0]),
                              "dCD": abs(ctx[usg.MIXED][2][1] - ctx[usg.FP64][2][1]),
                              "res_hist_rel": float(np.max(np.abs(ctx[usg.MIXED][3] - ctx[usg.FP64][3]) /
                                                           ctx[usg.FP64][3]))},
               "cl_fp64": ctx[usg.FP64][2][0], "cd_fp64": ctx[usg.FP64][2][1],
               "res_ratio_fp64": float(ctx[usg.FP64][3][-1] / ctx[usg.FP64][3][0]),
               "cell_it_s_fp64": ctx[usg.FP64][4], "cell_it_s_mixed": ctx[usg.MIXED][4]}
        # same state: the mixed context takes the FP64 run's final state
        sf, sm = ctx[usg.FP64][0], ctx[usg.MIXED][0]
        U, T = sf.U, sf.field(usg.F_TURB)
        sm.set_prim(U, T)
        sf.residual()
        sm.residual()
        ra = np.concatenate([sf.R, sf.field(usg.F_TRES)], axis=1)
        rb = np.concatenate([sm.R, sm.field(usg.F_TRES)], axis=1)
        mx, rms = _rel(rb - ra, ra)
        cla, cda = sf.forces()
        clb, cdb = sm.forces()
        row["same_state"] = {"residual_max_rel": dict(zip(EQS, mx)), "residual_rms_rel": dict(zip(EQS, rms)),
                             "dCL": abs(clb - cla), "dCD": abs(cdb - cda)}
        rows.append(row)
        sf.close()
        sm.close()
        print(f"# aoa {aoa}: trajectory max rel {row['trajectory']['max_rel']}; same-state residual "
              f"rms rel {row['same_state']['residual_rms_rel']}", file=sys.stderr, flush=True)
    return {"mesh_cells": m.nc, "iterations": iters, "order": order, "mach": mach, "reynolds": re, "cases": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=150)
    ap.add_argument("--aoa", type=float, nargs="+", default=[0.0, 2.0, 4.0, 6.0, 8.0, 10.0])
    ap.add_argument("--order", type=int, default=2, choices=[1, 2],
                    help="reconstruction order of both contexts (1: first order, which converges SST here)")
    a = ap.parse_args()
    print(json.dumps(run(a.iters, a.aoa, order=a.order)), flush=True)


if __name__ == "__main__":
    main()
