"""Overset connectivity (usg_overset_*, the serial restatement of UniTioga's
tioga::performConnectivity) on the cylinder-in-box pair: an O-grid of
256 x 96 quads around a unit cylinder (first spacing 1e-3, ratio 1.08,
capped at 0.05) over a 400 x 200 Cartesian background on [-10,30]x[-10,10].

The reference's connectivity needs MPI and HDF5 and cannot run here, so its
node iblanks and donors are not compared with reference output.  What the
tests pin instead:

* every donor stencil contains its receptor and reproduces linear fields
  (weights sum to 1, sum w x_v = x_p), as computeNodalWeight2D defines them;
* no cell holds both a hole node and a field node: one fringe layer of
  receptors separates them (nfringe = 1);
* the overset boundary's cell layer is receptors (mandatory: its nodes get
  resolution 1e15, tagBoundary), the cylinder's interior is hole;
* donors are finer than their receptors unless the receptor is mandatory;
* the result is symmetric under y -> -y, as the configuration is.
"""
import functools

import numpy as np
import pytest

from tmp_amd import mesh, usg


@functools.lru_cache(maxsize=None)
def c2():
    bg = mesh.background()
    og, flags = mesh.cylinder_ogrid()
    return og, bg, flags, usg.overset_connect([og, bg], [flags, None])


def test_ogrid_geometry():
    r = mesh.ogrid_radii(96, 0.5, 1e-3, 1.08, 0.05)
    assert r[1] - r[0] == pytest.approx(1e-3)
    assert r[2] - r[1] == pytest.approx(1.08e-3)
    assert np.max(np.diff(r)) == pytest.approx(0.05)
    assert 3.3 < r[-1] < 3.5
    og, flags = mesh.cylinder_ogrid()
    assert og.nc == 256 * 96 and flags.sum() == 256
    assert int((og.frc == mesh.WALL).sum()) == 256


def test_stencils_contain_and_reproduce_linear_fields():
    og, bg, _, res = c2()
    ms = [og, bg]
    for b, r in enumerate(res):
        assert len(r["recv_node"]) > 0
        xr = np.array([ms[rb].xy[rn] for rb, rn in zip(r["recv_block"], r["recv_node"])])
        xd = ms[b].xy[r["donor_node"]]
        w = r["w"]
        assert np.all(r["nw"] == 4)
        assert np.max(np.abs(w.sum(axis=1) - 1)) < 1e-14
        assert np.max(np.abs(np.einsum("nk,nkd->nd", w, xd) - xr)) < 1e-13
        assert w.min() > -1e-10 and w.max() < 1 + 1e-10
        # the donor cell is one of this block's cells and holds the vertices
        cn, cv = usg.cell_nodes(ms[b])
        for k in range(0, len(w), 37):
            assert set(r["donor_node"][k].tolist()) == set(cn[r["donor_cell"][k], :cv[r["donor_cell"][k]]].tolist())
        # receptors of this block's donors are nodes of the other block
        assert np.all(r["recv_block"] == 1 - b)


def test_receptor_iblank_matches_receptor_lists():
    og, bg, _, res = c2()
    for b in range(2):
        recv = set(res[1 - b]["recv_node"].tolist())
        nib = res[b]["node_iblank"]
        assert recv == set(np.nonzero(nib == -1)[0].tolist())


@pytest.mark.parametrize("b", [0, 1])
def test_fringe_separates_holes_from_field(b):
    og, bg, _, res = c2()
    m = [og, bg][b]
    cn, cv = usg.cell_nodes(m)
    nib = res[b]["node_iblank"]
    cib = res[b]["cell_iblank"]
    for c in range(m.nc):
        v = nib[cn[c, :cv[c]]]
        assert not ((v == 0).any() and (v == 1).any()), c
        # getCellIblanks: 0 with a hole vertex, -1 when every vertex receives
        want = 0 if (v == 0).any() else (-1 if (v == -1).all() else 1)
        assert cib[c] == want


def test_hole_and_mandatory_receptors():
    og, bg, _, res = c2()
    r_bg = np.hypot(bg.xy[:, 0], bg.xy[:, 1])
    nib = res[1]["node_iblank"]
    assert np.all(nib[r_bg < 0.5] == 0)              # inside the cylinder
    assert np.all(nib[r_bg > 3.5] == 1)              # beyond the O-grid
    assert (nib == 0).sum() > 1000 and (nib == -1).sum() > 100
    # the O-grid's outer cell layer: both node rings receive
    ni, nj = 256, 96
    oib = res[0]["node_iblank"].reshape(ni, nj + 1)
    assert np.all(oib[:, -2:] == -1)
    assert np.all(oib[:, :-2] == 1)
    assert np.all(res[0]["node_res"].reshape(ni, nj + 1)[:, -2:] == 1e15)


def test_donor_exclusion_depth():
    """Cells within mexclude = 3 layers of the overset boundary's cell layer
    are no donors (resolution 1e15): the four outer layers of the O-grid."""
    og, bg, _, res = c2()
    cres = res[0]["cell_res"]
    # O-grid cell (i, j) is cell i * 96 + j before renumbering
    layer = np.arange(og.nc) % 96
    assert np.all(cres[layer >= 92] == 1e15)
    assert np.all(cres[layer < 92] < 1e15)
    assert np.all(res[1]["cell_res"] < 1e15)       # the background has no overset boundary


def test_donors_are_finer_or_receptors_mandatory():
    og, bg, _, res = c2()
    for b in range(2):
        r = res[b]
        rec_res = res[1 - b]["node_res"][r["recv_node"]]
        don_res = r["cell_res"][r["donor_cell"]]
        assert np.all((don_res < rec_res) | (rec_res == 1e15))
        assert np.all(don_res < 1e15)


def test_mirror_symmetry():
    og, bg, _, res = c2()
    nib = res[1]["node_iblank"]
    key = {(round(x, 9), round(y, 9)): i for i, (x, y) in enumerate(bg.xy)}
    mirror = np.array([key[(round(x, 9), round(-y, 9))] for x, y in bg.xy])
    assert np.array_equal(nib, nib[mirror])


def test_blocks_without_overlap():
    a = mesh.rect(8, 8)
    b = mesh.rect(8, 8)
    b.xy = b.xy + np.array([5.0, 0.0])
    res = usg.overset_connect([a, b])
    for r in res:
        assert np.all(r["node_iblank"] == 1) and len(r["recv_node"]) == 0
