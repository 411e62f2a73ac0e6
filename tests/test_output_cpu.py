"""Solution outputs of the library (host only): the Tecplot cell-data file of
writeCellData (TopologyHolderStrategyOutput.cpp:466-554) and the wall C_p of
outPutCp (Euler2D.cpp:392-623).

The reference's writers need SAMRAI, MPI and HDF5 and cannot run here, so the
expected files below are derived by hand from the reference's source on a
two-cell mesh: its vertex order follows Brick4Cobalt2D::insert
(CellBrick.cpp:13-136) traced face by face, and its values are chosen so that
every number has a short exact form.
"""
import os

import numpy as np
import pytest

from tmp_amd import mesh, usg
from oracle import oracle as O

from .common import two_cells

HERE = os.path.dirname(os.path.abspath(__file__))


# rho, p, u, v per cell; free stream (rho, p, u, v) = (1, 1, 2, 0)
U2 = np.array([[1.0, 2.0, 3.0, 1e-5], [123456789.0, 0.5, 0.0, 1.0]])
FS2 = np.array([1.0, 1.0, 2.0, 0.0])


def test_cell_vertices_follow_the_reference_order():
    cnode, cnv = usg.cell_nodes(two_cells())
    # cell 0: (0,1) then (1,2) fills slot 2, (2,3) slot 3; cell 1: (1,2),
    # then (1,4) fills slot 3, (4,2) closes it and slot 3 moves to slot 2
    assert cnv.tolist() == [4, 3]
    assert cnode.tolist() == [[0, 1, 2, 3], [1, 2, 4, -1]]


@pytest.mark.parametrize("name", ["part2", "hybrid"])
def test_cell_vertices_match_the_oracle(name):
    if name == "part2":
        m = mesh.read_grd(os.path.join(HERE, "golden", "part2.grd"), -2, -2147483647)
    else:
        m = mesh.airfoil_hybrid(ni=96, dy1=2e-3, rfar=20.0, growth=1.25, seed=5, jitter=0.1)
    cnode, cnv = usg.cell_nodes(m)
    om = O.OracleMesh(m)
    assert np.array_equal(cnv, om.cnv)
    assert np.array_equal(cnode, om.cnode)


def test_write_cell_data_text(tmp_path):
    path = usg.write_cell_data(str(tmp_path / "flow"), two_cells(), U2, FS2, proc=2, mesh_tag=0)
    assert path.endswith("flow_1_3")
    want = ('TITLE ="Tioga output"\n'
            'VARIABLES="X","Y","Density","Pressure","U","V",\n'
            '"Cp","Ma"ZONE T="VOL_MIXED",N=5 E=2 ET = QUADRILATERAL, F = FEBLOCK\n'
            'VARLOCATION =  (1=NODAL,2=NODAL,3=CELLCENTERED,4=CELLCENTERED,5=CELLCENTERED,'
            '6=CELLCENTERED,7=CELLCENTERED,8=CELLCENTERED)\n'
            # X, Y of the nodes
            '0\n1\n1\n0\n2\n' '0\n0\n1\n1\n0.5\n'
            # density, pressure, u, v of the cells (6 significant digits)
            '1\n1.23457e+08\n' '2\n0.5\n' '3\n0\n' '1e-05\n1\n'
            # Cp = 2 (p - 1) / 4
            '0.5\n-0.25\n'
            # |u| / sqrt(p / rho): sqrt(9 + 1e-10) / sqrt(2), sqrt(2 * 123456789)
            '2.12132\n15713.5\n'
            # 1-based vertices, the triangle's last vertex repeated
            '1\t2\t3\t4\t\n' '2\t3\t5\t5\t\n')
    with open(path) as f:
        assert f.read() == want


def test_surface_cp_two_cells(tmp_path):
    table, curve = usg.surface_cp(two_cells(), U2, FS2)
    # wall nodes 0, 3 (face (3,0), normal (-1, 0)) and 4, 2 (face (4,2),
    # normal y > 0: negative counts); node C_p = (mean p - 1) * 2 / 4
    want = [[0, 0, 1, 0.5], [0, 1, 1, 0.5], [2, 0.5, -1, -0.25], [1, 1, -2, 0.125]]
    assert table.tolist() == want
    # nodes 0 and 3 share x: merged into the first
    assert curve.tolist() == [want[0], want[2], want[3]]
    usg.write_cp(str(tmp_path / "Cp_7"), curve)
    with open(tmp_path / "Cp_7") as f:
        assert f.read() == ('Title = "Dussin_exp"\nVARIABLES = "X","Cp"\nZONE T="Dussin_exp", I = 4, F=POINT\n'
                            '0\t0.5\n2\t-0.25\n1\t0.125\n0\t0.5\n')


def test_surface_cp_merge_takes_the_larger_count():
    # the triangle's wall face (4,2) replaced by a wall at x = 1 + 5e-5:
    # nodes 2 (x 1) and 4 (x 1.00005) merge, the larger |count| wins
    m = two_cells()
    m.xy[4] = [1.00005, 0.5]
    m.xy[2] = [1.0, 1.0]
    table, curve = usg.surface_cp(m, U2, FS2)
    # negative group in descending x: node 4 (count -1, Cp -0.25), then node
    # 2 (count -2, Cp 0.125) which replaces its C_p and count
    assert table[2].tolist() == [1.00005, 0.5, -1, -0.25]
    assert table[3].tolist() == [1, 1, -2, 0.125]
    assert curve[1].tolist() == [1.00005, 0.5, -2, 0.125]
    assert curve.shape[0] == 2


def test_surface_cp_airfoil_properties():
    m = mesh.airfoil_hybrid(ni=96, dy1=2e-3, rfar=20.0, growth=1.25)
    fs = usg.freestream_state(0.5, 0.0)
    rng = np.random.default_rng(3)
    U = np.column_stack([np.ones(m.nc), 1 + 0.1 * rng.random(m.nc), np.zeros(m.nc), np.zeros(m.nc)])
    table, curve = usg.surface_cp(m, U, fs)
    cnode, cnv = usg.cell_nodes(m)
    wall = np.unique(m.fnode[m.frc == mesh.WALL])
    assert table.shape[0] == wall.size
    # |count| = cells around the node; C_p from their mean pressure
    cells_of = {int(v): [] for v in wall}
    for c in range(m.nc):
        for v in cnode[c, :cnv[c]]:
            if int(v) in cells_of:
                cells_of[int(v)].append(c)
    by_xy = {(float(m.xy[v, 0]), float(m.xy[v, 1])): v for v in wall}
    for x, y, cnt, cp in table:
        cs = cells_of[by_xy[(x, y)]]
        assert abs(cnt) == len(cs)
        assert cp == pytest.approx((np.mean(U[cs, 1]) - fs[1]) * 2 / (fs[0] * (fs[2] ** 2 + fs[3] ** 2)),
                                   rel=1e-12, abs=1e-12)
    # upper surface (normals into the body point down) ascending in x, then
    # the lower surface descending: one loop around the section; the blunt
    # trailing-edge base (normal along x) may flag its upper corner too
    pos = table[:, 2] > 0
    body = table[:, 0] < 0.99
    assert np.all(table[pos & body, 1] >= 0) and np.all(table[~pos & body, 1] <= 0)
    assert np.all(np.diff(table[pos, 0]) >= 0) and np.all(np.diff(table[~pos, 0]) <= 0)
    assert pos.sum() > 10 and (~pos).sum() > 10
    # uniform free-stream pressure: C_p = 0 everywhere
    U[:, 1] = fs[1]
    _, c0 = usg.surface_cp(m, U, fs)
    assert np.all(c0[:, 3] == 0)


def test_freestream_state_matches_the_oracle():
    for mach, aoa in [(0.5, 2.0), (0.15, 10.0), (0.8, -3.0)]:
        fs = usg.freestream_state(mach, aoa)
        orc = O.OracleSolver(O.OracleMesh(two_cells()), O.EULER, mach, aoa, 1e6)
        assert np.array_equal(fs, np.array(list(orc.s.fs_primVar)[:4]))


def test_output_argument_errors(tmp_path):
    with pytest.raises(usg.UsgError):
        usg.write_cell_data(str(tmp_path / "x"), two_cells(), U2[:1], FS2)
    with pytest.raises(usg.UsgError):
        usg.write_cp(str(tmp_path / "c"), np.zeros((0, 4)))
    with pytest.raises(usg.UsgError):
        usg.write_cell_data(str(tmp_path / "missing" / "x"), two_cells(), U2, FS2)
