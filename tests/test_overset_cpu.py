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


@pytest.mark.parametrize("n", list(range(1, 40)) + [101, 256, 1000])
def test_median_partition(n):
    """The reference's own median test (UniTioga/mul_test.cpp:32-95): after
    find_median_wrap the split element (n/2 - 1 for even n, n/2 for odd n,
    the ADT's nleft - 1) has no larger element before it and no smaller one
    after it; the companion indices follow the values.  The partition keeps
    this for every even size.  For odd sizes the reference's routine misses
    it in a few percent of random inputs (its test prints such failures
    rather than stopping), and so does this restatement: there only the
    permutation is checked, plus a miss rate below 10 %."""
    rng = np.random.default_rng(n)
    misses = 0
    trials = 40
    for trial in range(trials):
        x0 = rng.uniform(100, 200, size=n)
        if trial % 4 == 3:
            x0 = np.round(x0 / 10) * 10          # ties
        x, ix = usg.overset_median(x0)
        assert np.array_equal(x, x0[ix])
        assert sorted(ix.tolist()) == list(range(n))
        imid = n // 2 - 1 if n % 2 == 0 else n // 2
        ok = bool(np.all(x[:imid] <= x[imid]) and np.all(x[imid + 1:] >= x[imid]))
        if n % 2 == 0:
            assert ok, trial
        misses += not ok
    assert misses <= 0.1 * trials


def test_oriented_box():
    """findOBB (utils.cpp:43-292): Jacobi eigenvectors of the node
    covariance (kaiser.cpp), then the box of the projections, 1.01 x the
    half extents.  findOBB copies the eigenvector matrix (eigenvectors in
    its columns) row by row into the box axes, so the axes are the principal
    directions reflected: a valid bounding box, tight only for axis-aligned
    clouds.  The restatement keeps that."""
    # the reference's kaiser_test cloud (mul_test.cpp:427-451): y = 2x + U(1, 2)
    rng = np.random.default_rng(7)
    xs = rng.uniform(10, 20, 1000)
    pts = np.column_stack([xs, 2 * xs + rng.uniform(1, 2, 1000)])
    xc, dxc, vec = usg.overset_obb(pts)
    assert np.allclose(vec @ vec.T, np.eye(2), atol=1e-12)
    principal = np.linalg.eigh(np.cov(pts.T))[1][:, 1]          # the (1, 2) direction
    assert abs(abs(principal @ np.array([1, 2])) / np.sqrt(5) - 1) < 1e-3
    assert max(abs(vec[:, 0] @ principal), abs(vec[:, 1] @ principal)) > 1 - 1e-9   # a column, not a row
    d = (pts - xc) @ vec.T
    assert np.all(np.abs(d) <= dxc * (1 + 1e-12))
    # axis-aligned rectangle: half sizes 1.01 x the half extents, centred
    u, v = np.meshgrid(np.linspace(-3, 3, 13), np.linspace(-1, 1, 5))
    pts = np.column_stack([u.ravel(), v.ravel()]) + np.array([5.0, -2.0])
    xc, dxc, vec = usg.overset_obb(pts)
    assert np.allclose(np.sort(dxc)[::-1], [3 * 1.01, 1 * 1.01], rtol=1e-12)
    assert np.allclose(np.abs(vec), np.eye(2)) or np.allclose(np.abs(vec), np.eye(2)[::-1])
    assert np.allclose(xc, [5.0, -2.0], atol=1e-12)
    # one and two points: the reference's special cases
    xc, dxc, vec = usg.overset_obb([[1.0, 2.0]])
    assert np.array_equal(dxc, [1e-3, 1e-3]) and np.array_equal(vec, np.eye(2))
    xc, dxc, vec = usg.overset_obb([[0.0, 0.0], [4.0, 0.0005]])
    # TIOGA_MAX(1e-3, d) * 0.5 expands to (1e-3 > d ? 1e-3 : d * 0.5)
    assert np.array_equal(dxc, [2.0, 1e-3])


# The reference's own overset scenario: part1.grd and part2.grd moved by
# (-0.7, 0.15) (Orchestra::init, Orchestra.cpp:204), wall tag -2, overset
# tag -2147483647, mexclude 1 and nfringe 2 (test.input, read at
# Orchestra.cpp:105-110).
@functools.lru_cache(maxsize=None)
def ref_pair():
    import os
    g = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    p1 = mesh.read_grd(os.path.join(g, "part1.grd"), -2, -2147483647)
    p2 = mesh.read_grd(os.path.join(g, "part2.grd"), -2, -2147483647)
    p2.xy = p2.xy + np.array([-0.7, 0.15])
    flags = [(m.frc == mesh.FARFIELD).astype(np.int32) for m in (p1, p2)]
    return (p1, p2), usg.overset_connect([p1, p2], flags, mexclude=1, nfringe=2)


def _inside(poly, pts):
    """Even-odd point-in-polygon test."""
    x, y = pts[:, 0], pts[:, 1]
    inside = np.zeros(len(pts), bool)
    n = len(poly)
    for i in range(n):
        (x0, y0), (x1, y1) = poly[i], poly[(i + 1) % n]
        cross = ((y0 > y) != (y1 > y)) & (x < (x1 - x0) * (y - y0) / np.where(y1 == y0, 1, y1 - y0) + x0)
        inside ^= cross
    return inside


def _wall_loop(m, tag=mesh.WALL):
    """The nodes of the faces tagged `tag` as one closed loop."""
    w = m.fnode[m.frc == tag]
    nxt = {int(a): int(b) for a, b in w}
    start = int(w[0, 0])
    loop, v = [start], nxt[start]
    while v != start:
        loop.append(v)
        v = nxt[v]
    assert len(loop) == len(w)
    return m.xy[loop]


def test_reference_pair_stencils_and_fringes():
    ms, res = ref_pair()
    for b, r in enumerate(res):
        assert len(r["recv_node"]) > 0
        xr = ms[1 - b].xy[r["recv_node"]]
        w = r["w"]
        xd = ms[b].xy[r["donor_node"]]
        assert np.max(np.abs(w.sum(axis=1) - 1)) < 1e-14
        assert np.max(np.abs(np.einsum("nk,nkd->nd", w, xd) - xr)) < 1e-13
        assert w.min() > -1e-10 and w.max() < 1 + 1e-10
        assert set(r["recv_node"].tolist()) == set(np.nonzero(res[1 - b]["node_iblank"] == -1)[0].tolist())
    for b in range(2):
        cn, cv = usg.cell_nodes(ms[b])
        nib = res[b]["node_iblank"]
        for c in range(ms[b].nc):
            v = nib[cn[c, :cv[c]]]
            assert not ((v == 0).any() and (v == 1).any())


def test_reference_pair_holes_inside_the_other_wall():
    (p1, p2), res = ref_pair()
    nib = res[0]["node_iblank"]
    inside = _inside(_wall_loop(p2), p1.xy)
    assert inside.sum() > 10
    assert np.all(nib[inside] == 0)
    # part2's wall is not cut by part1: no hole nodes in part2
    assert np.all(res[1]["node_iblank"] != 0)


def test_reference_pair_exclusion_depth():
    """mexclude = 1: the overset-boundary cell layer and one more layer of
    part2 may not donate; its overset-boundary nodes are mandatory
    receptors wherever part1 covers them (the others have no donor at all
    and stay field nodes, orphans of this two-mesh scenario)."""
    (p1, p2), res = ref_pair()
    # cell layers of part2 by face adjacency from its overset boundary
    nb = p2.neighbours()
    layer = np.full(p2.nc, -1)
    front = np.unique(p2.flc[p2.frc == mesh.FARFIELD])
    layer[front] = 0
    k = 0
    while front.size:
        nxt = np.unique(nb[front][nb[front] >= 0])
        nxt = nxt[layer[nxt] < 0]
        k += 1
        layer[nxt] = k
        front = nxt
    cres = res[1]["cell_res"]
    assert np.all(cres[layer <= 1] == 1e15)
    assert np.all(cres[layer >= 3] < 1e15)
    obc_nodes = np.unique(p2.fnode[p2.frc == mesh.FARFIELD])
    covered = _inside(_wall_loop(p1, mesh.FARFIELD), p2.xy[obc_nodes]) & ~_inside(_wall_loop(p1), p2.xy[obc_nodes])
    assert covered.sum() > 10 and (~covered).sum() > 10
    nib = res[1]["node_iblank"][obc_nodes]
    assert np.all(nib[~covered] == 1)
    assert np.all(nib != 0)
    # a covered node stays a field node only when every part1 cell around it
    # is itself excluded as a donor (within mexclude of part1's own boundary)
    cn, cv = usg.cell_nodes(p1)
    xv = p1.xy[np.where(cn >= 0, cn, cn[:, :1])]
    for v in obc_nodes[covered & (nib == 1)]:
        w = usg.nodal_weights(cv, xv, np.tile(p2.xy[v], (p1.nc, 1)))
        holds = np.all((w > -1e-10) & (w < 1 + 1e-10), axis=1)
        assert holds.any() and np.all(res[0]["cell_res"][holds] == 1e15)
    assert (nib[covered] == -1).sum() > 20
