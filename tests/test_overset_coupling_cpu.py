"""Cell-centred overset coupling and dual time stepping without a GPU.

* The product's composite stencil (usg_overset_cell_stencil) against the
  oracle's stage-by-stage restatement (oracle/overset_oracle.py) on the
  cylinder O-grid over the Cartesian background: the same blanking and
  receptor cells (integers, exactly), and the same receptor states for a
  smooth field to rounding.
* Closed forms: the stencil's weights sum to one and its donors are field
  cells; a linear field is reproduced exactly where the donor mesh is the
  uniform background (the node means of a uniform grid and the vertex mean
  of a cell are exact for linear fields).
* The same on the reference's own pair (part1 + part2, test.input).
* The dual-time term of the oracle's residual: with the free stream as the
  current state (zero spatial residual) and earlier time levels on a
  straight line in time, W^n = W - a dt and W^(n-1) = W - 2 a dt, the
  residual is V a for BDF2 (3W - 4W^n + W^(n-1)) / (2 dt) and for BDF1
  (W - W^n) / dt, and the LU-SGS diagonal gains c0 V.
"""
import functools

import numpy as np
import pytest

from oracle import oracle as O
from oracle.overset_oracle import OversetCoupling
from tmp_amd import mesh, usg


@functools.lru_cache(maxsize=None)
def c2():
    og, flags = mesh.cylinder_ogrid()
    bg = mesh.background()
    conn, st = usg.overset_coupling([og, bg], [flags, None])
    return og, bg, conn, st


def test_stencil_equals_the_staged_oracle():
    og, bg, conn, st = c2()
    oms = [O.OracleMesh(og), O.OracleMesh(bg)]
    cp = OversetCoupling(oms, conn)
    assert np.array_equal(np.concatenate(cp.blank), st["blank"])
    assert np.array_equal(np.concatenate([cp.recv[0], cp.recv[1] + og.nc]), st["recv"])
    assert (st["blank"][:og.nc] == -1).sum() == 256          # the O-grid's outer cell layer
    assert (st["blank"][og.nc:] == 0).sum() > 2000           # the hole in the background
    cc = [oms[0].cc.copy(), oms[1].cc.copy()]
    U = [np.column_stack([1 + 0.1 * np.sin(c[:, 0]), 1 + 0.1 * np.cos(c[:, 1]), 0.2 + 0.05 * c[:, 0],
                          0.01 * c[:, 1]]) for c in cc]
    res = cp.receptor_states(U)
    staged = np.concatenate([res[0][1], res[1][1]])
    Uall = np.concatenate(U)
    comp = np.stack([np.add.reduceat(st["w"] * Uall[st["donor"], j], st["off"][:-1]) for j in range(4)], 1)
    assert np.max(np.abs(comp - staged)) <= 1e-14 * np.max(np.abs(staged))


@functools.lru_cache(maxsize=None)
def ref_pair():
    """The reference's own overset scenario: part1.grd and part2.grd moved by
    (-0.7, 0.15) (Orchestra.cpp:204), wall tag -2, overset tag -2147483647,
    mexclude 1, nfringe 2 (test.input, read at Orchestra.cpp:105-110)."""
    import os
    g = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    p1 = mesh.read_grd(os.path.join(g, "part1.grd"), -2, -2147483647)
    p2 = mesh.read_grd(os.path.join(g, "part2.grd"), -2, -2147483647)
    p2.xy = p2.xy + np.array([-0.7, 0.15])
    flags = [(m.frc == mesh.FARFIELD).astype(np.int32) for m in (p1, p2)]
    conn, st = usg.overset_coupling([p1, p2], flags, mexclude=1, nfringe=2)
    return p1, p2, conn, st


def test_reference_pair_stencil_equals_the_staged_oracle():
    """On the reference's part1 + part2 pair: the same blanking and receptor
    cells as the oracle's staged coupling, the same receptor states for a
    smooth field, donors that are field cells and weights that sum to one.
    Both meshes receive (part1 around part2's element, part2 on its overset
    boundary where part1 covers it)."""
    p1, p2, conn, st = ref_pair()
    oms = [O.OracleMesh(p1), O.OracleMesh(p2)]
    cp = OversetCoupling(oms, conn)
    assert np.array_equal(np.concatenate(cp.blank), st["blank"])
    assert np.array_equal(np.concatenate([cp.recv[0], cp.recv[1] + p1.nc]), st["recv"])
    n1 = int((st["recv"] < p1.nc).sum())
    assert n1 > 0 and len(st["recv"]) - n1 > 0
    assert (st["blank"][:p1.nc] == 0).sum() > 0                 # part1 is cut by part2's wall
    assert np.all(st["blank"][st["donor"]] == 1)
    assert np.allclose(np.add.reduceat(st["w"], st["off"][:-1]), 1.0, rtol=0, atol=4e-16)
    cc = [oms[0].cc.copy(), oms[1].cc.copy()]
    U = [np.column_stack([1 + 0.1 * np.sin(3 * c[:, 0]), 1 + 0.1 * np.cos(2 * c[:, 1]), 0.3 + 0.05 * c[:, 0],
                          0.02 * c[:, 1]]) for c in cc]
    res = cp.receptor_states(U)
    staged = np.concatenate([res[0][1], res[1][1]])
    Uall = np.concatenate(U)
    comp = np.stack([np.add.reduceat(st["w"] * Uall[st["donor"], j], st["off"][:-1]) for j in range(4)], 1)
    assert np.max(np.abs(comp - staged)) <= 1e-14 * np.max(np.abs(staged))


def test_stencil_weights_donors_and_linear_fields():
    og, bg, conn, st = c2()
    off, donor, w = st["off"], st["donor"], st["w"]
    assert np.all(np.diff(off) > 0)
    assert np.allclose(np.add.reduceat(w, off[:-1]), 1.0, rtol=0, atol=4e-16)
    assert np.all(st["blank"][donor] == 1)
    assert np.all(st["blank"][st["recv"]] == -1)
    cc = np.concatenate([usg.mesh_geometry(og)["cc"], usg.mesh_geometry(bg)["cc"]])
    f = 0.3 + 1.7 * cc[:, 0] - 0.9 * cc[:, 1]
    val = np.add.reduceat(w * f[donor], off[:-1])
    on_ogrid = st["recv"] < og.nc            # interpolated from the uniform background
    assert np.max(np.abs(val - f[st["recv"]])[on_ogrid]) < 1e-13
    # from the stretched O-grid the node means are first order: still close
    assert np.max(np.abs(val - f[st["recv"]])[~on_ogrid]) < 2e-3


@pytest.mark.parametrize("order", [1, 2])
def test_dual_time_term_closed_form(order):
    m = mesh.rect(6, 5, jitter=0.15, seed=2)
    om = O.OracleMesh(m)
    s = O.OracleSolver(om, O.LAMINAR, 0.4, 10.0, 500.0)
    dt = 0.25
    a = np.array([0.01, -0.03, 0.02, 0.005])
    s.dual_time(dt, order)
    s.dual_time_advance()
    if order == 2:
        s.dual_time_advance()                      # BDF2 from the second physical step
    W = s.field("W", 4)
    s.field("Wt1", 4)[:] = W - a * dt
    s.field("Wt2", 4)[:] = W - 2 * a * dt
    s.residual()
    expect = om.vol[:, None] * a[None, :]
    assert np.allclose(s.R, expect, rtol=1e-10, atol=1e-13 * np.abs(expect).max())
    s.sweeps()
    c0 = 1.5 / dt if order == 2 else 1.0 / dt
    lamc, lamv = s.field("lamc", 1), s.field("lamv", 1)
    assert np.allclose(s.field("D", 4)[:, 0], lamc / 400.0 + 0.75 * lamc + lamv + c0 * om.vol, rtol=1e-13, atol=0)


def test_overset_cells_frozen_in_the_sweeps():
    """Hole and receptor cells take no increment and keep their state in an
    LU-SGS iteration and reconstruct first order (limiter 0, so a
    receptor's gradient, which reaches into hole cells, never enters a face
    state); the norm sums field cells only."""
    og, bg, conn, st = c2()
    s = O.OracleSolver(O.OracleMesh(bg), O.EULER, 0.3, 0.0, 1e6)
    blank = st["blank"][og.nc:]
    U0 = s.U.copy()
    U0[:, 0] *= 1.0 + 0.01 * np.sin(np.arange(bg.nc))
    s.set_prim(U0)
    s.set_blank(blank)
    s.lusgs_step()
    fz = blank != 1
    ph = s.field("phi", 4)
    assert np.all(ph[fz] == 0.0) and np.all(ph[~fz] > 0.0)
    assert np.all(s.field("dW", 4)[fz] == 0.0) and np.array_equal(s.U[fz], U0[fz])
    assert np.any(s.U[~fz] != U0[~fz])
    r = s.R[:, 0] / s.mesh.vol
    assert abs(s.res_l2() - np.sqrt(np.sum(r[~fz] ** 2))) <= 1e-12 * s.res_l2()
