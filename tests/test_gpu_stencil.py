"""Donor-stencil apply kernel (usg_stencil_apply) against the oracle
(orc_stencil_apply, the restatement of getInterpolatedSolution +
updateSolnData): bit-identical in both layouts, bad entries skipped, and
linear fields reproduced through weights from usg_nodal_weights on a real
mesh."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tmp_amd import usg

from .common import part2

pytestmark = pytest.mark.gpu


def _dev(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _stencil(rng, npts_src, npts_dst, nrecv):
    recv = rng.choice(npts_dst, size=nrecv, replace=False).astype(np.int32)
    nw = rng.choice([3, 4], size=nrecv).astype(np.int32)
    donor = rng.integers(0, npts_src, size=(nrecv, 4)).astype(np.int32)
    w = rng.uniform(0, 1, size=(nrecv, 4))
    w[nw == 3, 3] = 0.0
    w /= w.sum(axis=1, keepdims=True)
    return recv, nw, donor, w


@pytest.mark.parametrize("layout", [usg.ROW, usg.COLUMN], ids=["row", "column"])
def test_apply_matches_oracle(layout):
    rng = np.random.default_rng(21)
    ns, nd, nvar, nrecv = 5000, 3000, 6, 700
    src = rng.normal(size=(ns, nvar))
    dst0 = rng.normal(size=(nd, nvar))
    recv, nw, donor, w = _stencil(rng, ns, nd, nrecv)
    ref = O.stencil_apply(src, dst0.copy(), recv, nw, donor.reshape(-1), w.reshape(-1))
    s = _dev(src if layout == usg.ROW else src.T, torch.float64)
    d = _dev(dst0 if layout == usg.ROW else dst0.T, torch.float64)
    keep = [_dev(recv, torch.int32), _dev(nw, torch.int32), _dev(donor, torch.int32), _dev(w, torch.float64)]
    torch.cuda.synchronize()
    usg.stencil_apply(s.data_ptr(), d.data_ptr(), nvar, layout, ns, nd, nrecv, *(k.data_ptr() for k in keep))
    torch.cuda.synchronize()
    out = d.cpu().numpy()
    out = out if layout == usg.ROW else out.T
    assert np.array_equal(out, ref)


def test_out_of_range_entries_are_skipped():
    rng = np.random.default_rng(22)
    ns, nd, nvar = 100, 50, 4
    src = rng.normal(size=(ns, nvar))
    recv = np.array([3, 4, 999], np.int32)                   # receptor 999 out of range
    nw = np.array([2, 3, 1], np.int32)
    donor = np.array([[1, 2, 0, 0], [5, ns + 7, 6, 0], [0, 0, 0, 0]], np.int32)  # one donor out of range
    w = np.array([[0.5, 0.5, 0, 0], [0.2, 0.3, 0.5, 0], [1, 0, 0, 0]])
    d = torch.zeros((nd, nvar), dtype=torch.float64, device="cuda")
    keep = [_dev(recv, torch.int32), _dev(nw, torch.int32), _dev(donor, torch.int32), _dev(w, torch.float64)]
    usg.stencil_apply(_dev(src, torch.float64).data_ptr(), d.data_ptr(), nvar, usg.ROW, ns, nd, 3,
                      *(k.data_ptr() for k in keep))
    torch.cuda.synchronize()
    out = d.cpu().numpy()
    assert np.array_equal(out[3], 0.5 * src[1] + 0.5 * src[2])
    assert np.array_equal(out[4], (0.0 + 0.2 * src[5]) + 0.5 * src[6])
    assert np.count_nonzero(out[[i for i in range(nd) if i not in (3, 4)]]) == 0


def test_linear_field_reproduced_on_mesh_cells():
    """Receptor points inside cells of the part2 fixture, weights from the
    product's usg_nodal_weights on the cell vertices, a linear nodal field:
    the apply returns the field at the points."""
    m = part2()
    g = usg.mesh_geometry(m)
    rng = np.random.default_rng(23)
    cells = rng.choice(m.nc, size=400, replace=False)
    # vertices of each cell from its faces
    verts = [[] for _ in range(m.nc)]
    for f in range(m.nf):
        for c in (m.flc[f], m.frc[f]):
            if c >= 0:
                for nd in m.fnode[f]:
                    if nd not in verts[c]:
                        verts[c].append(nd)
    nv = np.array([len(verts[c]) for c in cells], np.int32)
    donor = np.zeros((len(cells), 4), np.int32)
    xv = np.zeros((len(cells), 4, 2))
    for r, c in enumerate(cells):
        vs = verts[c]
        donor[r, :len(vs)] = vs
        xv[r, :len(vs)] = m.xy[vs]
    xp = g["cc"][cells] + 0.1 * (xv[:, 0] - g["cc"][cells])
    w = usg.nodal_weights(nv, xv, xp)
    a = np.array([0.3, -1.2, 2.0, 0.7])
    b = np.array([1.0, 0.5, -0.25, 3.0])
    c = np.array([-2.0, 0.1, 1.5, 0.0])
    q = a + np.outer(m.xy[:, 0], b) + np.outer(m.xy[:, 1], c)
    d = torch.zeros((len(cells), 4), dtype=torch.float64, device="cuda")
    recv = np.arange(len(cells), dtype=np.int32)
    keep = [_dev(recv, torch.int32), _dev(nv, torch.int32), _dev(donor, torch.int32), _dev(w, torch.float64)]
    usg.stencil_apply(_dev(q, torch.float64).data_ptr(), d.data_ptr(), 4, usg.ROW, m.nn, len(cells), len(cells),
                      *(k.data_ptr() for k in keep))
    torch.cuda.synchronize()
    expect = a + np.outer(xp[:, 0], b) + np.outer(xp[:, 1], c)
    assert np.max(np.abs(d.cpu().numpy() - expect)) <= 1e-11 * np.max(np.abs(expect))


def test_overset_receptors_interpolate_on_gpu():
    """The cylinder-in-box connectivity (usg_overset_*) drives the stencil
    kernel: node values of one block interpolated to the other block's
    receptor nodes match the oracle bit for bit and reproduce a linear field
    to rounding."""
    from tmp_amd import mesh
    bg = mesh.background()
    og, flags = mesh.cylinder_ogrid()
    res = usg.overset_connect([og, bg], [flags, None])
    ms = [og, bg]
    for b in range(2):
        r = res[b]
        src_m, dst_m = ms[b], ms[1 - b]
        n = len(r["recv_node"])
        lin = np.array([0.5, 2.0, -1.5])
        src = np.stack([lin[0] + lin[1] * src_m.xy[:, 0] + lin[2] * src_m.xy[:, 1],
                        np.sin(src_m.xy[:, 0]) * np.cos(src_m.xy[:, 1])], axis=1)
        dst0 = np.zeros((dst_m.nn, 2))
        ref = O.stencil_apply(src, dst0.copy(), r["recv_node"], r["nw"], r["donor_node"].reshape(-1),
                              r["w"].reshape(-1))
        s = _dev(src.T, torch.float64)
        d = _dev(dst0.T, torch.float64)
        keep = [_dev(r["recv_node"], torch.int32), _dev(r["nw"], torch.int32), _dev(r["donor_node"], torch.int32),
                _dev(r["w"], torch.float64)]
        torch.cuda.synchronize()
        usg.stencil_apply(s.data_ptr(), d.data_ptr(), 2, usg.COLUMN, src_m.nn, dst_m.nn, n,
                          *(k.data_ptr() for k in keep))
        torch.cuda.synchronize()
        out = d.cpu().numpy().T
        assert np.array_equal(out, ref)
        xr = dst_m.xy[r["recv_node"]]
        exact = lin[0] + lin[1] * xr[:, 0] + lin[2] * xr[:, 1]
        assert np.max(np.abs(out[r["recv_node"], 0] - exact)) < 1e-12


def test_tioga_adapter_driver():
    """The C++ Tioga adapter (tmp_amd/host/usg_strategy.h) driven the way
    Orchestra::init drives tioga (Orchestra.cpp:200-264) on the reference's
    part1 + part2 pair: after dataUpdate every receptor node holds the other
    mesh's value (node values i+1 per mesh), every other node its own, and
    the iblank counts equal the Python front end's."""
    import os
    import subprocess
    from tmp_amd import mesh
    from .common import GOLDEN
    exe = os.path.join(os.path.dirname(usg.LIB_PATH), "usg_run")
    out = subprocess.run([exe, "--overset", os.path.join(GOLDEN, "part1.grd"), os.path.join(GOLDEN, "part2.grd"),
                          "-2", "-2147483647", "-0.7", "0.15", "1", "2"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [dict(zip(l.split()[2::2], map(int, l.split()[3::2]))) for l in out.stdout.splitlines()
            if l.startswith("mesh")]
    assert len(rows) == 2
    p1 = mesh.read_grd(os.path.join(GOLDEN, "part1.grd"), -2, -2147483647)
    p2 = mesh.read_grd(os.path.join(GOLDEN, "part2.grd"), -2, -2147483647)
    p2.xy = p2.xy + np.array([-0.7, 0.15])
    res = usg.overset_connect([p1, p2], [(m.frc == mesh.FARFIELD).astype(np.int32) for m in (p1, p2)],
                              mexclude=1, nfringe=2)
    for r, py in zip(rows, res):
        nib = py["node_iblank"]
        assert r["field"] == (nib == 1).sum() and r["hole"] == (nib == 0).sum() and r["receptor"] == (nib == -1).sum()
        assert r["receptor"] > 0
        assert r["took_other"] == r["receptor"]
        assert r["unchanged"] == r["field"] + r["hole"]
        assert r["cell_receptors"] == (py["cell_iblank"] == -1).sum()
