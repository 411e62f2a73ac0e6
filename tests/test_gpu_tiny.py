"""The smallest meshes through the C ABI, against the CPU FP64 oracle.

* One triangle and one quadrilateral, each with a wall face and far-field
  faces: no interior face, seven of the eight spatial parts empty, and both
  LU-SGS sweeps a single wave that runs inside the cell kernel.
* Two cells (a quadrilateral and a triangle sharing one face): two waves,
  one launch of each sweep.

Bars as in test_gpu_parity.py: 1e-13 on one residual evaluation, 1e-12 on
trajectories.  A batch mixing these cases with a normal mesh must equal the
single-case runs bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tmp_amd import usg

from .common import one_quad, one_triangle, part2, perturbed, relerr, two_cells

pytestmark = pytest.mark.gpu

TOL_RES = 1e-13
TOL_TRAJ = 1e-12

MESHES = [("one-triangle", one_triangle), ("one-quad", one_quad), ("two-cells", two_cells)]
FLOWS = [("euler", O.EULER, 0.5, 2.0, 1e6), ("laminar", O.LAMINAR, 0.5, 2.0, 5000.0),
         ("sst", O.SST, 0.15, 4.0, 6e6)]
GRID = [(mn, mk, fn, fl, ma, ao, re) for mn, mk in MESHES for fn, fl, ma, ao, re in FLOWS]
IDS = [f"{g[0]}-{g[2]}" for g in GRID]


@pytest.mark.parametrize("mname,mk,fname,flow,mach,aoa,re", GRID, ids=IDS)
def test_tiny_residual(mname, mk, fname, flow, mach, aoa, re):
    m = mk()
    geom = usg.mesh_geometry(m)
    orc = O.OracleSolver(O.OracleMesh(m), flow, mach, aoa, re)
    gpu = usg.Solver(m, flow=flow, mach=mach, aoa=aoa, reynolds=re)
    U = perturbed(m, geom["cc"], list(orc.s.fs_primVar)[:4], seed=3)
    orc.set_prim(U)
    gpu.set_prim(U)
    orc.residual()
    gpu.residual()
    checks = {"R": (gpu.field(usg.F_RES), orc.R), "grad": (gpu.field(usg.F_GRAD), orc.field("gradU", 8)),
              "lamc": (gpu.field(usg.F_LAMC), orc.field("lamc", 1)), "phi": (gpu.field(usg.F_PHI), orc.field("phi", 4))}
    if flow != O.EULER:
        checks["lamv"] = (gpu.field(usg.F_LAMV), orc.field("lamv", 1))
    if flow == O.SST:
        checks["tR"] = (gpu.field(usg.F_TRES), orc.field("tR", 2))
    for k, (g, c) in checks.items():
        e = np.max(relerr(g, c))
        print(f"{mname} {fname} {k}: rel {e:.3e}")
        assert np.all(np.isfinite(g)), k
        assert e <= TOL_RES, (k, e)


@pytest.mark.parametrize("mname,mk,fname,flow,mach,aoa,re", GRID, ids=IDS)
def test_tiny_lusgs(mname, mk, fname, flow, mach, aoa, re):
    m = mk()
    orc = O.OracleSolver(O.OracleMesh(m), flow, mach, aoa, re)
    gpu = usg.Solver(m, flow=flow, mach=mach, aoa=aoa, reynolds=re)
    assert gpu.waves() == ((1, 1) if m.nc == 1 else (2, 2))
    hist_o = []
    for _ in range(10):
        orc.lusgs_step()
        hist_o.append(orc.res_l2())
    hist_g = gpu.lusgs(10, history=True)
    eU = np.max(relerr(gpu.U, orc.U))
    eh = np.max(np.abs(hist_g - np.array(hist_o)) / np.array(hist_o))
    print(f"{mname} {fname}: U rel {eU:.3e} history rel {eh:.3e}")
    assert np.all(np.isfinite(gpu.U))
    assert eU <= TOL_TRAJ and eh <= TOL_TRAJ
    if flow == O.SST:
        assert np.max(relerr(gpu.field(usg.F_TURB), orc.field("tU", 2))) <= TOL_TRAJ
    # the oracle's force reads the wall states of its last residual
    # evaluation: evaluate both at the final state first
    orc.residual()
    gpu.residual()
    clo, cdo = orc.forces()
    clg, cdg = gpu.forces()
    assert abs(clg - clo) <= 1e-12 * max(1.0, abs(clo)) and abs(cdg - cdo) <= 1e-12 * max(1.0, abs(cdo))


@pytest.mark.parametrize("mname,mk", MESHES, ids=[x[0] for x in MESHES])
def test_tiny_rk5(mname, mk):
    m = mk()
    orc = O.OracleSolver(O.OracleMesh(m), O.EULER, 0.5, 2.0, 1e6)
    gpu = usg.Solver(m, flow=usg.EULER, mach=0.5, aoa=2.0)
    for _ in range(3):
        orc.rk5_step()
    gpu.rk5(3)
    e = np.max(relerr(gpu.U, orc.U))
    print(f"{mname} rk5: U rel {e:.3e}")
    assert e <= TOL_TRAJ


def test_batch_with_tiny_cases():
    """One-cell and two-cell cases beside a 2222-cell one in one batch: the
    batch's parts and waves then mix empty and full slices."""
    meshes = [one_triangle(), part2(), two_cells(), one_quad()]
    machs, aoas = [0.5, 0.5, 0.4, 0.6], [2.0, 3.0, 1.0, -1.0]
    b = usg.BatchSolver(meshes, flow=usg.EULER, mach=machs, aoa=aoas)
    singles = [usg.Solver(m, flow=usg.EULER, mach=machs[i], aoa=aoas[i]) for i, m in enumerate(meshes)]
    hb = b.lusgs(7, history=True)
    for i, s in enumerate(singles):
        hs = s.lusgs(7, history=True)
        assert np.array_equal(b.U[b.case_slice(i)], s.U), i
        assert np.array_equal(hb[:, i], hs), i
