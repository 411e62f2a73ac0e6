"""The C ABI (include/usg.h) without a GPU: libusg.so loads, exports every
function the header declares, the Python front end binds exactly that set,
and argument errors come back as status codes before any device work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from tmp_amd import usg

from .common import one_quad, one_triangle, part2, two_cells

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")


def declared(name=None):
    """Functions declared by include/<name> (every header when None)."""
    names = set()
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h") or (name and h != name):
            continue
        text = re.sub(r"/\*.*?\*/", "", open(os.path.join(INCLUDE, h)).read(), flags=re.S)
        names |= set(re.findall(r"^[a-z][\w\s\*]*?\b(usg_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_parses():
    names = declared()
    assert "usg_ctx_create" in names and "usg_last_error" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    L = usg.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binds_exactly_the_headers():
    from tmp_amd import mlp
    assert sorted(usg.EXPORTS) == declared("usg.h")
    assert sorted(mlp.EXPORTS) == declared("usg_mlp.h")


def _params(flow=usg.EULER, precision=usg.FP64):
    return usg._Params(flow, precision, 0.5, 0.0, 1e6, 400.0, 0.8)


def test_bad_flow_rejected_before_device():
    m = part2()
    d, keep = usg._desc(m)
    h = C.c_void_p()
    rc = usg.lib().usg_ctx_create(C.byref(d), C.byref(_params(flow=7)), C.byref(h))
    assert rc != 0 and not h.value
    assert b"Undefined Flow Type" in usg.lib().usg_last_error()


def test_null_arguments_rejected():
    L = usg.lib()
    assert L.usg_ctx_create(None, None, None) != 0
    assert L.usg_lusgs_step(None, 1, None) != 0
    assert L.usg_residual(None) != 0


def test_mesh_build_rejects_bad_tags():
    m = part2()
    frc = m.frc.copy()
    frc[np.argmin(frc)] = -5          # not a cell, not WALL/FARFIELD
    bad = usg._mesh.Mesh(m.xy, m.fnode, m.flc, frc, m.nc, "bad")
    with pytest.raises(usg.UsgError):
        usg.mesh_geometry(bad)


def test_batch_rejects_mixed_flow_types():
    m = part2()
    d, keep = usg._desc(m)
    descs = (usg._MeshDesc * 2)(d, d)
    params = (usg._Params * 2)(_params(usg.EULER), _params(usg.LAMINAR))
    h = C.c_void_p()
    assert usg.lib().usg_batch_create(descs, params, 2, C.byref(h)) != 0 and not h.value
    assert b"share flow type" in usg.lib().usg_last_error()
    assert usg.lib().usg_batch_create(descs, params, 0, C.byref(h)) != 0


def test_empty_mesh_rejected():
    z = np.zeros(0, np.int32)
    empty = usg._mesh.Mesh(np.zeros((0, 2)), np.zeros((0, 2), np.int32), z, z, 0, "empty")
    with pytest.raises(usg.UsgError, match="empty mesh"):
        usg.mesh_geometry(empty)


def test_tiny_mesh_geometry():
    """One- and two-cell meshes: shoelace volume, vertex-mean centre, the
    projected wall distance |(x_c - x_f) . n_f| and one wave per cell
    colour (UnstructBlock.h:265-298, cellVolume.cpp:6-78)."""
    for m, nw in ((one_triangle(), 1), (one_quad(), 1), (two_cells(), 2)):
        g = usg.mesh_geometry(m)
        assert (g["nfwd"], g["nbwd"]) == (nw, nw)
        cnode, cnv = usg.cell_nodes(m)
        for c in range(m.nc):
            v = m.xy[cnode[c, :cnv[c]]]
            x, y = v[:, 0], v[:, 1]
            area = 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))
            assert g["vol"][c] == pytest.approx(area, rel=1e-15)
            assert np.allclose(g["cc"][c], v.mean(axis=0), rtol=0, atol=1e-15)
            walls = [f for f in range(m.nf) if m.frc[f] == -2]
            d = min(abs(np.dot(g["cc"][c] - g["fc"][f], g["fn"][f])) for f in walls)
            assert g["wdist"][c] == pytest.approx(d, rel=1e-14)


def test_oversized_mesh_rejected_before_arrays_are_read():
    """8 * nc must fit the kernels' 32-bit offsets: a descriptor claiming
    more cells is refused before its (here far too short) arrays are read."""
    m = part2()
    d, keep = usg._desc(m)
    d.nc = 268435456
    h = C.c_void_p()
    assert usg.lib().usg_ctx_create(C.byref(d), C.byref(_params()), C.byref(h)) == -1   # USG_ERR_ARG
    assert b"32-bit array offsets" in usg.lib().usg_last_error() and not h.value
    d.nc = 150000000
    descs = (usg._MeshDesc * 2)(d, d)
    params = (usg._Params * 2)(_params(), _params())
    assert usg.lib().usg_batch_create(descs, params, 2, C.byref(h)) == -1
    assert b"32-bit array offsets" in usg.lib().usg_last_error()
