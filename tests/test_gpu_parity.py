"""GPU (gfx950 HIP, through the C ABI) against the CPU FP64 oracle.

Bar: FP64 results within 1e-12 relative of the oracle (max |gpu - cpu| /
max |cpu| per variable).  Single residual evaluations are held tighter
(1e-13); the only non-bit-exact operations are libm ``pow``/``tanh`` (glibc
on the CPU, ROCm device libm on the GPU, 1-2 ulp apart) at far-field faces,
in Sutherland's law, in the HLLC Roe sound speed and in SST F1/F2.  Level-
scheduled sweeps against the oracle's sequential sweeps is the LU-SGS check.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tmp_amd import usg

from .common import airfoil_small, bump, delaunay_small, part2, perturbed, relerr

pytestmark = pytest.mark.gpu

TOL_RES = 1e-13
TOL_TRAJ = 1e-12


def _pair(m, flow, mach, aoa, re, seed=0, perturb=True):
    geom = usg.mesh_geometry(m)
    om = O.OracleMesh(m)
    orc = O.OracleSolver(om, flow, mach, aoa, re)
    gpu = usg.Solver(m, flow=flow, mach=mach, aoa=aoa, reynolds=re)
    if perturb:
        U = perturbed(m, geom["cc"], list(orc.s.fs_primVar)[:4], seed=seed)
        orc.set_prim(U)
        gpu.set_prim(U)
    return orc, gpu, om


CASES = [
    ("part2", part2, O.EULER, 0.5, 2.0, 1e6),
    ("bump", bump, O.EULER, 0.5, 0.0, 1e6),
    ("airfoil-laminar", airfoil_small, O.LAMINAR, 0.5, 2.0, 5000.0),
    ("airfoil-sst", airfoil_small, O.SST, 0.15, 4.0, 6e6),
    ("delaunay-sst", delaunay_small, O.SST, 0.15, 0.0, 6e6),
]


@pytest.mark.parametrize("name,mk,flow,mach,aoa,re", CASES, ids=[c[0] for c in CASES])
def test_residual_parity(name, mk, flow, mach, aoa, re):
    m = mk()
    orc, gpu, om = _pair(m, flow, mach, aoa, re)
    orc.residual()
    gpu.residual()
    checks = {
        "R": (gpu.field(usg.F_RES), orc.R),
        "grad": (gpu.field(usg.F_GRAD), orc.field("gradU", 8)),
        "lamc": (gpu.field(usg.F_LAMC), orc.field("lamc", 1)),
        "phi": (gpu.field(usg.F_PHI), orc.field("phi", 4)),
    }
    if flow != O.EULER:
        checks["lamv"] = (gpu.field(usg.F_LAMV), orc.field("lamv", 1))
        checks["gradT"] = (gpu.field(usg.F_GRADT), orc.field("gradT", 2))
    if flow == O.SST:
        checks["tR"] = (gpu.field(usg.F_TRES), orc.field("tR", 2))
        checks["F1"] = (gpu.field(usg.F_F1), orc.field("F1", 1))
    for k, (g, c) in checks.items():
        e = np.max(relerr(g, c))
        print(f"{name} {k}: rel {e:.3e} exact {np.mean(g == c):.4f}")
        assert np.all(np.isfinite(g)), k
        assert e <= TOL_RES, (k, e)


@pytest.mark.parametrize("name,mk,flow,mach,aoa,re,n", [
    ("bump-euler", bump, O.EULER, 0.5, 0.0, 1e6, 40),
    ("part2-euler", part2, O.EULER, 0.5, 2.0, 1e6, 40),
    ("airfoil-laminar", airfoil_small, O.LAMINAR, 0.5, 2.0, 5000.0, 30),
    ("airfoil-sst", airfoil_small, O.SST, 0.15, 4.0, 6e6, 20),
    ("delaunay-laminar", delaunay_small, O.LAMINAR, 0.2, 0.0, 100.0, 30),
    ("delaunay-sst", delaunay_small, O.SST, 0.15, 0.0, 6e6, 20),
], ids=["bump-euler", "part2-euler", "airfoil-laminar", "airfoil-sst", "delaunay-laminar", "delaunay-sst"])
def test_lusgs_trajectory(name, mk, flow, mach, aoa, re, n):
    """n LU-SGS iterations from the free stream: GPU wave-scheduled sweeps vs
    the oracle's sequential sweeps (LUSGS_Integrator.cpp:492-561)."""
    m = mk()
    orc, gpu, om = _pair(m, flow, mach, aoa, re, perturb=False)
    hist_o = []
    for _ in range(n):
        orc.lusgs_step()
        hist_o.append(orc.res_l2())
    hist_g = gpu.lusgs(n, history=True)
    eU = np.max(relerr(gpu.U, orc.U))
    eh = np.max(np.abs(hist_g - np.array(hist_o)) / np.array(hist_o))
    print(f"{name}: U rel {eU:.3e}  residual-history rel {eh:.3e}  waves {gpu.waves()}")
    assert np.all(np.isfinite(gpu.U))
    assert eU <= TOL_TRAJ
    assert eh <= TOL_TRAJ
    # the reference's printed monitor, sum of dW_rho^2 of the last iteration
    # (LUSGS_Integrator.cpp:131-153)
    d = orc.field("dW", 4)[:, 0]
    mon = float(np.cumsum(d * d)[-1])
    em = abs(gpu.dw_rho_sq()[0] - mon) / mon
    print(f"{name}: sum dW_rho^2 {mon:.6e} rel {em:.3e}")
    assert em <= TOL_TRAJ
    if flow == O.SST:
        et = np.max(relerr(gpu.field(usg.F_TURB), orc.field("tU", 2)))
        print(f"{name}: (k, omega) rel {et:.3e}")
        assert et <= TOL_TRAJ


@pytest.mark.parametrize("name,mk,flow,mach,aoa,re", [CASES[1], CASES[3]], ids=["bump", "airfoil-sst"])
def test_sweep_increments_and_state_fields(name, mk, flow, mach, aoa, re):
    """One LU-SGS iteration from a perturbed state: every component of the
    forward and backward increments dW1, dW (and the SST tdW1, tdW), the
    conservative state W and the SST (rho k, rho omega) against the oracle's,
    through the field exports (LUSGS_Integrator.cpp:492-561)."""
    m = mk()
    orc, gpu, om = _pair(m, flow, mach, aoa, re)
    orc.lusgs_step()
    gpu.lusgs(1)
    checks = {"dW1": (usg.F_DW1, "dW1", 4), "dW": (usg.F_DW, "dW", 4), "W": (usg.F_CONS, "W", 4)}
    if flow == O.SST:
        checks.update({"tdW": (usg.F_TDW, "tdW", 2), "tW": (usg.F_TCONS, "tW", 2)})
    for k, (fid, on, w) in checks.items():
        g, c = gpu.field(fid), orc.field(on, w)
        e = np.max(relerr(g, c))
        print(f"{name} {k}: rel {e:.3e}")
        assert np.all(np.isfinite(g)) and e <= TOL_TRAJ, (k, e)


@pytest.mark.parametrize("name,mk,flow", [("part2-euler", part2, O.EULER),
                                          ("airfoil-laminar", airfoil_small, O.LAMINAR)],
                         ids=["part2-euler", "airfoil-laminar"])
def test_rk5_trajectory(name, mk, flow):
    """RK5 (RungeKuttaStrategy.h:39-72): 10 steps = 50 residual evaluations."""
    m = mk()
    re = 5000.0 if flow == O.LAMINAR else 1e6
    orc, gpu, om = _pair(m, flow, 0.5, 2.0, re, perturb=False)
    for _ in range(10):
        orc.rk5_step()
    gpu.rk5(10)
    e = np.max(relerr(gpu.U, orc.U))
    print(f"{name} rk5: U rel {e:.3e}")
    assert e <= TOL_TRAJ
    # the reference's RK5 monitor (RungeKuttaStrategy.h:87-105)
    mon = float(np.cumsum(np.abs(orc.U[:, 0] - orc.field("W0", 4)[:, 0]))[-1])
    em = abs(gpu.rk_monitor()[0] - mon) / mon
    print(f"{name} rk5: sum |d rho| {mon:.6e} rel {em:.3e}")
    assert mon > 0 and em <= TOL_TRAJ


def test_deterministic():
    """Two runs of the same work are bit-identical (no FP atomics)."""
    m = bump()
    a = usg.Solver(m, flow=usg.EULER, mach=0.5)
    b = usg.Solver(m, flow=usg.EULER, mach=0.5)
    ha = a.lusgs(10, history=True)
    hb = b.lusgs(10, history=True)
    assert np.array_equal(a.U, b.U)
    assert np.array_equal(ha, hb)


@pytest.mark.parametrize("name,mk,flow,mach,aoa,re", [("part2-euler", part2, O.EULER, 0.5, 2.0, 1e6),
                                                     ("airfoil-laminar", airfoil_small, O.LAMINAR, 0.5, 2.0, 5000.0)],
                         ids=["part2-euler", "airfoil-laminar"])
def test_forces_match_oracle(name, mk, flow, mach, aoa, re):
    """Wall force: pressure p n A, minus the viscous flux tau . n A for
    laminar flow (tree-summed on the GPU, sequential in the oracle)."""
    m = mk()
    orc, gpu, om = _pair(m, flow, mach, aoa, re, perturb=False)
    for _ in range(5):
        orc.lusgs_step()
    gpu.lusgs(5)
    orc.residual()
    gpu.residual()
    clo, cdo = orc.forces()
    clg, cdg = gpu.forces()
    assert abs(clg - clo) <= 1e-12 * max(1.0, abs(clo))
    assert abs(cdg - cdo) <= 1e-12 * max(1.0, abs(cdo))


def test_cpp_adapter_driver():
    """The C++ adapters (tmp_amd/host/usg_strategy.h: UnstructReader,
    WrapSolver, GpuLUSGSIntegrator) driven like the reference's serial
    scenario reproduce the oracle."""
    import os
    import subprocess
    from .common import GOLDEN
    exe = os.path.join(os.path.dirname(usg.LIB_PATH), "usg_run")
    out = subprocess.run([exe, os.path.join(GOLDEN, "part2.grd"), "-2", "-2147483647", "Euler2D", "LUSGS",
                          "0.5", "1e6", "2.0", "10", "1e-30"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    line = [l for l in out.stdout.splitlines() if l.startswith("iterations")][0]
    tok = line.split()
    res = float(tok[3])
    u0 = [float(x.strip("(),")) for x in line.split("=")[1].split()]
    orc = O.OracleSolver(O.OracleMesh(part2()), O.EULER, 0.5, 2.0, 1e6)
    for _ in range(10):
        orc.lusgs_step()
    assert abs(res - orc.res_l2()) <= 1e-12 * orc.res_l2()
    assert np.allclose(u0, orc.U[0], rtol=1e-12, atol=0)


@pytest.mark.parametrize("flow", [usg.EULER, usg.SST])
def test_graph_replays_match_eager(flow):
    """Captured two-iteration graphs replayed with eager work and host
    synchronisation in between give bit-identical states and residual
    histories to eager launches, after every call (odd counts included)."""
    m = airfoil_small()
    kw = dict(flow=flow, mach=0.15 if flow == usg.SST else 0.5, aoa=3.0, reynolds=6e6)
    g = usg.Solver(m, **kw)
    e = usg.Solver(m, **kw)
    e.graphs(False)
    for n in (2, 4, 3, 2, 6, 1):
        hg = g.lusgs(n, history=True)
        he = e.lusgs(n, history=True)
        assert np.array_equal(hg, he)
        assert np.array_equal(g.U, e.U)
        g.residual()          # eager work between replays
        e.residual()
        assert np.array_equal(g.R, e.R)


@pytest.mark.parametrize("name,mk,flow,mach,aoa,re", CASES, ids=[c[0] for c in CASES])
def test_fused_face_cell_equals_two_launches(name, mk, flow, mach, aoa, re):
    """The fused face-cell launch (face record in LDS per tile) gives the
    two-launch path's residual fields and LU-SGS trajectory bit for bit."""
    m = mk()
    a = usg.Solver(m, flow=flow, mach=mach, aoa=aoa, reynolds=re)
    b = usg.Solver(m, flow=flow, mach=mach, aoa=aoa, reynolds=re)
    b.fused(False)
    U = perturbed(m, usg.mesh_geometry(m)["cc"], a.freestream(), seed=4)
    a.set_prim(U)
    b.set_prim(U)
    a.residual()
    b.residual()
    fields = [usg.F_RES, usg.F_LAMC, usg.F_DT] + ([usg.F_LAMV] if flow != usg.EULER else []) + \
             ([usg.F_TRES] if flow == usg.SST else [])
    for fid in fields:
        assert np.array_equal(a.field(fid), b.field(fid)), fid
    ha = a.lusgs(6, history=True)
    hb = b.lusgs(6, history=True)
    assert np.array_equal(ha, hb) and np.array_equal(a.U, b.U)
    assert a.forces() == b.forces()


@pytest.mark.parametrize("name,mk,flow,mach,aoa,re", CASES[:3], ids=[c[0] for c in CASES[:3]])
def test_face_states_match_oracle(name, mk, flow, mach, aoa, re):
    """The U_edge getter (Euler2D::updateEdgeValues' face states) of a
    perturbed state equals the oracle's U_edge bit for bit (libm pow in the
    far-field state aside)."""
    m = mk()
    orc, gpu, om = _pair(m, flow, mach, aoa, re)
    orc.residual()
    e = np.max(relerr(gpu.field(usg.F_UEDGE), orc.field("U_edge", 4, "nf")))
    assert e <= TOL_RES, e


def test_sst_wall_distance_matches_oracle():
    """The SST model's wall distance (the distance to the nearest wall face
    segment, computed on the GPU at context creation) equals the oracle's
    bit for bit, on a single mesh and inside a batch of two."""
    m = airfoil_small()
    om = O.OracleMesh(m)
    gpu = usg.Solver(m, flow=usg.SST, mach=0.15, aoa=4.0, reynolds=6e6)
    assert np.array_equal(gpu.field(usg.F_WDIST), om.wdist_wall)
    p2 = part2()
    b = usg.BatchSolver([p2, m], flow=usg.SST, mach=0.15, aoa=[0.0, 4.0], reynolds=6e6)
    w = b.field(usg.F_WDIST)
    assert np.array_equal(w[b.case_slice(1)], om.wdist_wall)
    op2 = O.OracleMesh(p2)
    assert np.array_equal(w[b.case_slice(0)], op2.wdist_wall)


@pytest.mark.parametrize("name,mk,flow,mach,aoa,re", [CASES[1], CASES[3]], ids=["bump", "airfoil-sst"])
def test_first_order_trajectory(name, mk, flow, mach, aoa, re):
    """First-order reconstruction (usg_set_order(1)): 20 LU-SGS iterations
    against the oracle with the same order."""
    m = mk()
    orc, gpu, om = _pair(m, flow, mach, aoa, re, perturb=False)
    orc.set_order(1)
    gpu.set_order(1)
    for _ in range(20):
        orc.lusgs_step()
    gpu.lusgs(20)
    e = np.max(relerr(gpu.U, orc.U))
    print(f"{name} first order: U rel {e:.3e}")
    assert e <= TOL_TRAJ
