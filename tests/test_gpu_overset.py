"""Overset coupling and dual time stepping on the GPU, through the C ABI,
against the oracle (FP64 bar 1e-12 relative).

The C2 configuration: laminar flow (M 0.2, Re 100) past a unit cylinder,
an O-grid of 256x96 quads over a 400x200 Cartesian background on
[-10,30]x[-10,10], two blocks of one batch context, dual time stepping
(BDF1 then BDF2) with LU-SGS sub-iterations, the receptor cells set from
their donors at the start of every sub-iteration.  The oracle runs the two
blocks as two serial solvers coupled by its stage-by-stage restatement of
the interpolation (oracle/overset_oracle.py).
"""
import functools

import numpy as np
import pytest

from oracle import oracle as O
from oracle.overset_oracle import OversetCoupling
from tmp_amd import mesh, usg

from .common import airfoil_small, relerr

pytestmark = pytest.mark.gpu

TOL_TRAJ = 1e-12


@functools.lru_cache(maxsize=None)
def c2():
    og, flags = mesh.cylinder_ogrid()
    bg = mesh.background()
    conn, st = usg.overset_coupling([og, bg], [flags, None])
    return og, bg, conn, st


def _c2_pair(dt):
    og, bg, conn, st = c2()
    gpu = usg.BatchSolver([og, bg], flow=usg.LAMINAR, mach=0.2, aoa=0.0, reynolds=100.0)
    gpu.set_overset(st)
    gpu.dual_time(dt, 2)
    oms = [O.OracleMesh(og), O.OracleMesh(bg)]
    cp = OversetCoupling(oms, conn)
    orc = [O.OracleSolver(om, O.LAMINAR, 0.2, 0.0, 100.0) for om in oms]
    for s, b in zip(orc, cp.blank):
        s.set_blank(b)
        s.dual_time(dt, 2)
    return gpu, orc, cp, oms


def test_c2_overset_dual_time_against_oracle():
    dt, nstep, nsub = 0.5, 3, 8
    gpu, orc, cp, oms = _c2_pair(dt)
    og, bg, conn, st = c2()
    for n in range(nstep):
        gpu.dual_time_advance()
        for s in orc:
            s.dual_time_advance()
        ho = []
        for _ in range(nsub):
            cp.apply(orc)
            for s in orc:
                s.lusgs_step()
            ho.append([s.res_l2() for s in orc])
        hg = gpu.lusgs(nsub, history=True)
        ho = np.array(ho)
        U = gpu.U
        # relative to the state's magnitude (a per-variable norm would divide by
        # the background's v, which is round-off while the free stream holds)
        eU = max(np.max(np.abs(U[gpu.case_slice(b)] - orc[b].U)) / np.max(np.abs(orc[b].U)) for b in range(2))
        # the background's norm starts at round-off (free stream) until the
        # cylinder's disturbance reaches it: compare against the larger norm
        eh = np.max(np.abs(hg - ho) / np.max(ho, axis=1, keepdims=True))
        print(f"C2 step {n}: U rel {eU:.3e}, history rel {eh:.3e}, O-grid residual {ho[0, 0]:.3e} -> "
              f"{ho[-1, 0]:.3e}", flush=True)
        assert np.all(np.isfinite(U))
        assert eU <= TOL_TRAJ and eh <= TOL_TRAJ
    clg, cdg = gpu.forces()
    for s in orc:
        s.residual()
    clo, cdo = orc[0].forces()
    print(f"C2 cylinder C_L {clg[0]:.6e} (oracle {clo:.6e}), C_D {cdg[0]:.6e} (oracle {cdo:.6e})", flush=True)
    assert abs(clg[0] - clo) <= 1e-12 * max(1.0, abs(clo)) and abs(cdg[0] - cdo) <= 1e-12 * abs(cdo)


def test_c2_blanked_cells_keep_their_state():
    """Hole cells never change; receptor cells hold exactly their donors'
    interpolation of the state at the start of the last iteration."""
    gpu, orc, cp, oms = _c2_pair(0.5)
    og, bg, conn, st = c2()
    U0 = gpu.U
    gpu.dual_time_advance()
    gpu.lusgs(3)
    U = gpu.U
    hole = st["blank"] == 0
    assert np.array_equal(U[hole], U0[hole])
    assert np.any(U[st["blank"] == 1] != U0[st["blank"] == 1])
    gpu.residual()
    ph = gpu.field(usg.F_PHI)
    assert np.all(ph[st["blank"] != 1] == 0.0) and np.all(ph[st["blank"] == 1] > 0.0)


def test_dual_time_single_mesh_against_oracle():
    """Dual time on one laminar airfoil mesh (no overset): 3 physical steps
    of 5 sub-iterations each, BDF1 then BDF2."""
    m = airfoil_small()
    dt = 0.05
    gpu = usg.Solver(m, flow=usg.LAMINAR, mach=0.5, aoa=2.0, reynolds=5000.0)
    gpu.dual_time(dt, 2)
    orc = O.OracleSolver(O.OracleMesh(m), O.LAMINAR, 0.5, 2.0, 5000.0)
    orc.dual_time(dt, 2)
    for n in range(3):
        gpu.dual_time_advance()
        orc.dual_time_advance()
        ho = []
        for _ in range(5):
            orc.lusgs_step()
            ho.append(orc.res_l2())
        hg = gpu.lusgs(5, history=True)
        eU = np.max(relerr(gpu.U, orc.U))
        eh = np.max(np.abs(hg - np.array(ho)) / np.array(ho))
        print(f"dual time step {n}: U rel {eU:.3e}, history rel {eh:.3e}", flush=True)
        assert eU <= TOL_TRAJ and eh <= TOL_TRAJ


def test_overset_dual_time_graph_replays_match_eager():
    """With the coupling and dual time on, captured two-iteration graphs
    give bit-identical states and histories to eager launches across
    physical steps (the BDF coefficients change between the first and the
    second step, which rebuilds the graphs)."""
    g, _, _, _ = _c2_pair(0.5)
    e, _, _, _ = _c2_pair(0.5)
    e.graphs(False)
    for n in range(3):
        g.dual_time_advance()
        e.dual_time_advance()
        hg = g.lusgs(4, history=True)
        he = e.lusgs(4, history=True)
        assert np.array_equal(hg, he)
        assert np.array_equal(g.U, e.U)


def test_reference_pair_coupled_lusgs_against_oracle():
    """The reference's own overset scenario (part1 + part2 moved by
    (-0.7, 0.15), mexclude 1, nfringe 2; ViscousFlow2D at M 0.85, AoA 0,
    Re 2000, test.input) coupled in the iteration, steady LU-SGS: 15
    iterations, state and both meshes' residual histories against the
    oracle's two coupled solvers.  The AMR background that surrounds the
    pair in the reference is out of scope, so receptor cells without donors
    stay field cells with their far-field faces."""
    from .test_overset_coupling_cpu import ref_pair
    p1, p2, conn, st = ref_pair()
    gpu = usg.BatchSolver([p1, p2], flow=usg.LAMINAR, mach=0.85, aoa=0.0, reynolds=2000.0)
    gpu.set_overset(st)
    oms = [O.OracleMesh(p1), O.OracleMesh(p2)]
    cp = OversetCoupling(oms, conn)
    orc = [O.OracleSolver(om, O.LAMINAR, 0.85, 0.0, 2000.0) for om in oms]
    for s, b in zip(orc, cp.blank):
        s.set_blank(b)
    n = 15
    ho = []
    for _ in range(n):
        cp.apply(orc)
        for s in orc:
            s.lusgs_step()
        ho.append([s.res_l2() for s in orc])
    hg = gpu.lusgs(n, history=True)
    ho = np.array(ho)
    U = gpu.U
    eU = max(np.max(np.abs(U[gpu.case_slice(b)] - orc[b].U)) / np.max(np.abs(orc[b].U)) for b in range(2))
    eh = np.max(np.abs(hg - ho) / ho)
    print(f"reference pair: {n} coupled LU-SGS, U rel {eU:.3e}, history rel {eh:.3e}, part1 residual "
          f"{ho[0, 0]:.3e} -> {ho[-1, 0]:.3e}", flush=True)
    assert np.all(np.isfinite(U)) and eU <= TOL_TRAJ and eh <= TOL_TRAJ


def test_set_overset_rejects_bad_stencils():
    og, bg, conn, st = c2()
    gpu = usg.BatchSolver([og, bg], flow=usg.LAMINAR, mach=0.2, aoa=0.0, reynolds=100.0)
    bad = dict(st)
    bad["donor"] = st["donor"].copy()
    bad["donor"][0] = st["recv"][1]                  # a receptor as donor: would race
    with pytest.raises(usg.UsgError, match="not a field cell"):
        gpu.set_overset(bad)
    bad = dict(st)
    bad["recv"] = st["recv"].copy()
    bad["recv"][1] = bad["recv"][0]                   # the same receptor twice
    with pytest.raises(usg.UsgError, match="distinct receptor"):
        gpu.set_overset(bad)
    with pytest.raises(usg.UsgError, match="overset coupling and dual time"):
        gpu.set_overset(st)
        gpu.rk5(1)
