"""Batched many-case solves (usg_batch_create) on the GPU.

A batch is a block-diagonal union of independent cases, so every case must
reproduce a separate single-case context bit for bit: state, per-case
residual history and forces.  Single-case contexts are themselves pinned to
the CPU oracle by test_gpu_parity.py.  The residual history also pins the
two launch forms of the norm: a single case runs it as blocks appended to its
sweep launches, a batch as separate kernels (the same code on the same
cells).
"""
import numpy as np
import pytest

from tmp_amd import usg

from .common import airfoil_small, bump, part2

pytestmark = pytest.mark.gpu


def _check_cases(b, singles, hb, hs):
    for i, s in enumerate(singles):
        sl = b.case_slice(i)
        assert np.array_equal(b.U[sl], s.U), i
        assert np.array_equal(hb[:, i], hs[i]), i


@pytest.mark.parametrize("precision", [usg.FP64, usg.MIXED], ids=["fp64", "mixed"])
def test_batch_euler_equals_single_cases(precision):
    meshes = [bump(), airfoil_small(), part2()]
    machs, aoas = [0.5, 0.4, 0.5], [0.0, 2.0, 3.0]
    b = usg.BatchSolver(meshes, flow=usg.EULER, mach=machs, aoa=aoas, precision=precision)
    singles = [usg.Solver(m, flow=usg.EULER, mach=machs[i], aoa=aoas[i], precision=precision)
               for i, m in enumerate(meshes)]
    assert b.ncase == 3 and b.nc == sum(m.nc for m in meshes)
    for n in (6, 3):          # graph replays and an eager odd iteration
        hb = b.lusgs(n, history=True)
        hs = [s.lusgs(n, history=True) for s in singles]
        assert hb.shape == (n, 3)
        _check_cases(b, singles, hb, hs)
    cl, cd = b.forces()
    for i, s in enumerate(singles):
        assert (cl[i], cd[i]) == s.forces()
    r = b.res_l2()
    b.residual()
    for i, s in enumerate(singles):
        s.residual()
    assert np.array_equal(b.res_l2(), np.array([s.res_l2() for s in singles]))
    assert r.shape == (3,)


def test_batch_sst_equals_single_cases():
    m = airfoil_small()
    res = [6e6, 3e6]
    b = usg.BatchSolver([m, m], flow=usg.SST, mach=0.15, aoa=[4.0, 6.0], reynolds=res)
    singles = [usg.Solver(m, flow=usg.SST, mach=0.15, aoa=a, reynolds=r) for a, r in zip([4.0, 6.0], res)]
    hb = b.lusgs(8, history=True)
    hs = [s.lusgs(8, history=True) for s in singles]
    _check_cases(b, singles, hb, hs)
    tb = b.field(usg.F_TURB)
    for i, s in enumerate(singles):
        assert np.array_equal(tb[b.case_slice(i)], s.field(usg.F_TURB))


def test_batch_rk5_equals_single_cases():
    meshes = [part2(), airfoil_small()]
    b = usg.BatchSolver(meshes, flow=usg.LAMINAR, mach=[0.5, 0.3], aoa=[2.0, 1.0], reynolds=[5000.0, 8000.0])
    singles = [usg.Solver(m, flow=usg.LAMINAR, mach=ma, aoa=a, reynolds=r)
               for m, ma, a, r in zip(meshes, [0.5, 0.3], [2.0, 1.0], [5000.0, 8000.0])]
    hb = b.rk5(4, history=True)
    hs = [s.rk5(4, history=True) for s in singles]
    _check_cases(b, singles, hb, hs)



def test_autotune_restores_state_and_blocks_do_not_change_results():
    """Tuning leaves the state as it found it, and any block sizes and
    either face-cell launch scheme give the same bits (one thread per face
    or cell, whatever the block size)."""
    m = airfoil_small()
    a = usg.Solver(m, flow=usg.SST, mach=0.15, aoa=4.0, reynolds=6e6)
    b = usg.Solver(m, flow=usg.SST, mach=0.15, aoa=4.0, reynolds=6e6)
    a.lusgs(3)
    b.lusgs(3)
    best = a.autotune(2)
    assert all(t in (64, 128, 256) for t in best[:4]) and best[4] in (0, 1) and a.launch_config() == best
    assert np.array_equal(a.U, b.U)
    a.set_launch_config([64, 128, 64, 128, 1 - best[4]])
    ha = a.lusgs(5, history=True)
    hb = b.lusgs(5, history=True)
    assert np.array_equal(ha, hb) and np.array_equal(a.U, b.U)
    assert np.array_equal(a.field(usg.F_TURB), b.field(usg.F_TURB))


def test_failed_case_is_frozen_and_reported():
    """A case whose residual fails the HLLC p* check (a negative pressure
    makes the sound speed NaN) is frozen at its state and reported; the
    other cases run on unchanged.  A single context fails the call instead,
    with the reference's message."""
    m = part2()
    b = usg.BatchSolver([m, m], flow=usg.EULER, mach=0.5, aoa=2.0)
    U = b.U.copy()
    bad = U[b.case_slice(1)]
    bad[0, 1] = -1.0                      # pressure of one cell of case 1
    U[b.case_slice(1)] = bad
    b.set_prim(U)
    b.lusgs(4)
    assert b.status().tolist() == [0, 1]
    assert np.array_equal(b.U[b.case_slice(1)], bad)
    s = usg.Solver(m, flow=usg.EULER, mach=0.5, aoa=2.0)
    s.lusgs(4)
    assert np.array_equal(b.U[b.case_slice(0)], s.U)
    one = usg.Solver(m, flow=usg.EULER, mach=0.5, aoa=2.0)
    one.set_prim(bad)
    with pytest.raises(usg.UsgError, match="p star"):
        one.lusgs(1)
    b.set_prim(b.U * 0 + s.U[0])          # a fresh state clears the status
    assert b.status().tolist() == [0, 0]
