"""The colour-blocked numbering of tmp_amd.mesh.prepare (host only).

* The colouring is proper: no two neighbours share a colour.
* Three colours cover all but a few cells (a fourth only where no local
  repair removes a defect), so the LU-SGS sweeps have at most 4 waves.
* The numbering is a permutation grouped by colour, interior faces keep
  lC < rC with the normal out of the left cell, and the product's wave
  schedule then uses the colour blocks as waves.
"""
import numpy as np
import pytest

from tmp_amd import mesh, usg

from .common import one_quad, part2, two_cells

MESHES = [("airfoil", lambda: mesh.airfoil_hybrid(ni=64, dy1=1e-4, rfar=20.0)),
          ("bump", lambda: mesh.channel_bump(seed=1001)),
          ("rect-tri", lambda: mesh.rect(12, 10, lx=2.0, ly=1.5, jitter=0.15, seed=3, triangles=True)),
          ("part2", part2), ("two-cells", two_cells), ("one-quad", one_quad)]


def _colours(m):
    c = mesh.cell_centres(m)
    key = mesh._hilbert_keys(c[:, 0], c[:, 1])
    nbl = [[o for o in r if o >= 0] for r in m.neighbours().tolist()]
    return np.array(mesh._three_colours(nbl, np.argsort(key, kind="stable").tolist())), nbl


@pytest.mark.parametrize("name,mk", MESHES, ids=[x[0] for x in MESHES])
def test_colouring_is_proper_with_few_colours(name, mk):
    m = mk()
    col, nbl = _colours(m)
    assert col.min() == 0 and col.max() <= 3
    inner = m.frc >= 0
    assert not np.any(col[m.flc[inner]] == col[m.frc[inner]])
    sizes = np.bincount(col)
    if len(sizes) == 4:
        assert sizes[3] <= max(16, m.nc // 5000), sizes


@pytest.mark.parametrize("name,mk", MESHES, ids=[x[0] for x in MESHES])
def test_prepared_numbering(name, mk):
    m0 = mk()
    m, ncol = mesh.prepare(m0)
    assert m.nc == m0.nc and m.nf == m0.nf and ncol <= 4
    inner = m.frc >= 0
    assert np.all(m.flc[inner] < m.frc[inner])
    # same cells: the multiset of cell areas is unchanged by the renumbering
    # (to rounding: a cell's vertex order, and so the diagonal a quad's area
    # is split along, follows its faces' file order)
    a0 = np.sort(usg.mesh_geometry(m0)["vol"])
    g = usg.mesh_geometry(m)
    assert np.allclose(np.sort(g["vol"]), a0, rtol=1e-12, atol=0)
    # normals still point out of the left cell: from its centre to the face
    d = g["fc"][inner] - g["cc"][m.flc[inner]]
    assert np.all(np.sum(d * g["fn"][inner], axis=1) > 0)
    # colour blocks are the sweeps' waves
    assert g["nfwd"] == ncol and g["nbwd"] == ncol


def test_plain_greedy_would_need_five_colours():
    """Why the repair exists: greedy over the Hilbert order alone leaves
    cells whose neighbours hold all three colours."""
    m = mesh.airfoil_hybrid(ni=64, dy1=1e-4, rfar=20.0)
    c = mesh.cell_centres(m)
    order = np.argsort(mesh._hilbert_keys(c[:, 0], c[:, 1]), kind="stable").tolist()
    nbl = [[o for o in r if o >= 0] for r in m.neighbours().tolist()]
    col = [-1] * m.nc
    for v in order:
        used = {col[o] for o in nbl[v] if col[o] >= 0}
        col[v] = min(k for k in range(6) if k not in used)
    assert max(col) >= 3


@pytest.mark.parametrize("name,mk", MESHES, ids=[x[0] for x in MESHES])
def test_library_renumbering_equals_prepare(name, mk):
    """usg_mesh_renumber (C++) and mesh.prepare (Python) give the same
    numbering: same faces in the same order, same colour count; perm maps
    new cells to old ones (the cell areas follow it)."""
    m0 = mk()
    mp, ncp = mesh.prepare(m0)
    mc, perm, ncc = usg.mesh_renumber(m0)
    assert ncc == ncp
    assert np.array_equal(mc.fnode, mp.fnode)
    assert np.array_equal(mc.flc, mp.flc) and np.array_equal(mc.frc, mp.frc)
    assert np.array_equal(np.sort(perm), np.arange(m0.nc))
    v0 = usg.mesh_geometry(m0)["vol"]
    v1 = usg.mesh_geometry(mc)["vol"]
    assert np.allclose(v1, v0[perm], rtol=1e-12, atol=0)


def test_library_renumbering_rejects_bad_input():
    z = np.zeros(0, np.int32)
    with pytest.raises(usg.UsgError, match="empty mesh"):
        usg.mesh_renumber(mesh.Mesh(np.zeros((0, 2)), np.zeros((0, 2), np.int32), z, z, 0, "empty"))
    m = part2()
    bad = mesh.Mesh(m.xy, m.fnode, m.flc, np.where(m.frc >= 0, m.nc + 5, m.frc).astype(np.int32), m.nc, "bad")
    with pytest.raises(usg.UsgError, match="out of range"):
        usg.mesh_renumber(bad)
