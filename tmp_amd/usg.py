"""Python front end of libusg.so (the C ABI in include/usg.h).

Mirrors the reference's driver-facing interface: a :class:`Solver` is what
``WrapSolver`` holds after ``setTopology``/``setParams``/``init``
(UnstructSolver/SolverWrapper.h:22-68) -- one mesh, one flow strategy
(``Euler2D`` / ``ViscousFlow2D``, laminar or SST) and its integrator (LU-SGS or
RK5) -- running on the GPU.  Host arrays use the reference's AoS layout
(``U[cell, var]`` with ``U = (rho, p, u, v)``).

The library is loaded from the package directory; there is no CPU fallback:
if ``libusg.so`` is missing or no GPU is visible, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import mesh as _mesh

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libusg.so")

EULER, LAMINAR, SST = 0, 1, 2
FP64, MIXED = 0, 1
F_PRIM, F_CONS, F_RES, F_GRAD, F_LAMC, F_PHI, F_DW1, F_DW = range(8)
F_LAMV, F_GRADT, F_TURB, F_TRES, F_MUT, F_F1, F_DT, F_TDW, F_TCONS, F_W0 = range(8, 18)
F_UEDGE, F_WDIST = 18, 19
K_GRAD, K_FACE, K_CELL, K_SWEEP = range(4)

_FIELD_WIDTH = {F_PRIM: 4, F_CONS: 4, F_RES: 4, F_GRAD: 8, F_LAMC: 1, F_PHI: 4, F_DW1: 4, F_DW: 4,
                F_LAMV: 1, F_GRADT: 2, F_TURB: 2, F_TRES: 2, F_MUT: 1, F_F1: 1, F_DT: 1, F_TDW: 2, F_TCONS: 2, F_W0: 4,
                F_WDIST: 1}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class UsgError(RuntimeError):
    pass


class _MeshDesc(C.Structure):
    _fields_ = [("nn", C.c_int), ("nf", C.c_int), ("nc", C.c_int), ("xy", _dp), ("fnode", _ip),
                ("flc", _ip), ("frc", _ip)]


class _Params(C.Structure):
    _fields_ = [("flow", C.c_int), ("precision", C.c_int), ("mach", C.c_double), ("aoa_deg", C.c_double),
                ("reynolds", C.c_double), ("cfl_lusgs", C.c_double), ("cfl_rk", C.c_double)]


EXPORTS = ["usg_ctx_create", "usg_ctx_destroy", "usg_set_stream", "usg_sync", "usg_init_freestream",
           "usg_set_prim", "usg_residual", "usg_lusgs_step", "usg_rk5_step", "usg_get_field", "usg_res_l2",
           "usg_forces", "usg_sweep_waves", "usg_get_geometry", "usg_mesh_build", "usg_mesh_renumber", "usg_timing_enable",
           "usg_kernel_time", "usg_stencil_apply", "usg_copy_kernel", "usg_use_graphs", "usg_last_error",
           "usg_batch_create", "usg_case_count", "usg_set_block", "usg_get_block", "usg_autotune",
           "usg_device_arch", "usg_case_status", "usg_nodal_weights", "usg_mesh_cells", "usg_freestream_state",
           "usg_write_cell_data", "usg_surface_cp", "usg_write_cp", "usg_set_state", "usg_overset_create",
           "usg_overset_destroy", "usg_overset_iblank", "usg_overset_receptors", "usg_overset_resolution",
           "usg_overset_median", "usg_overset_obb", "usg_stencil_apply_host", "usg_stencil_apply_host_noalloc",
           "usg_overset_cell_stencil", "usg_set_overset", "usg_dual_time", "usg_dual_time_advance", "usg_set_cfl",
           "usg_use_fused", "usg_get_fused", "usg_set_order"]

_lib = None


def lib():
    """Load libusg.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise UsgError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.usg_ctx_create.argtypes = [C.POINTER(_MeshDesc), C.POINTER(_Params), C.POINTER(vp)]
        L.usg_ctx_destroy.argtypes = [vp]
        L.usg_set_stream.argtypes = [vp, vp]
        L.usg_sync.argtypes = [vp]
        L.usg_init_freestream.argtypes = [vp]
        L.usg_set_prim.argtypes = [vp, _dp, _dp]
        L.usg_set_state.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
        L.usg_residual.argtypes = [vp]
        L.usg_lusgs_step.argtypes = [vp, C.c_int, _dp]
        L.usg_rk5_step.argtypes = [vp, C.c_int, _dp]
        L.usg_get_field.argtypes = [vp, C.c_int, _dp, C.c_long]
        L.usg_res_l2.argtypes = [vp, _dp]
        L.usg_forces.argtypes = [vp, _dp, _dp]
        L.usg_sweep_waves.argtypes = [vp, _ip, _ip]
        L.usg_get_geometry.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.usg_mesh_build.argtypes = [C.POINTER(_MeshDesc), _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip]
        L.usg_timing_enable.argtypes = [vp, C.c_int]
        L.usg_kernel_time.argtypes = [vp, _dp, C.POINTER(C.c_long)]
        L.usg_stencil_apply.argtypes = [vp, vp, C.c_int, C.c_int, C.c_long, C.c_long, C.c_int, vp, vp, vp, vp, vp]
        L.usg_nodal_weights.argtypes = [C.c_int, _ip, _dp, _dp, _dp]
        L.usg_copy_kernel.argtypes = [vp, vp, C.c_long, vp]
        L.usg_use_graphs.argtypes = [vp, C.c_int]
        L.usg_batch_create.argtypes = [C.POINTER(_MeshDesc), C.POINTER(_Params), C.c_int, C.POINTER(vp)]
        L.usg_case_count.argtypes = [vp, _ip, _ip]
        L.usg_set_block.argtypes = [vp, C.c_int, C.c_int]
        L.usg_get_block.argtypes = [vp, C.c_int, _ip]
        L.usg_autotune.argtypes = [vp, C.c_int, _ip]
        L.usg_device_arch.argtypes = [C.c_char_p, C.c_int]
        L.usg_case_status.argtypes = [vp, _ip]
        L.usg_mesh_cells.argtypes = [C.POINTER(_MeshDesc), _ip, _ip]
        L.usg_freestream_state.argtypes = [C.c_double, C.c_double, _dp]
        L.usg_write_cell_data.argtypes = [C.POINTER(_MeshDesc), _dp, _dp, C.c_char_p, C.c_int, C.c_int]
        L.usg_surface_cp.argtypes = [C.POINTER(_MeshDesc), _dp, _dp, C.c_int, _dp, _dp, _ip, _ip]
        L.usg_write_cp.argtypes = [C.c_char_p, _dp, C.c_int]
        L.usg_overset_create.argtypes = [C.POINTER(_MeshDesc), C.POINTER(_ip), C.c_int, C.c_int, C.c_int,
                                         C.POINTER(vp)]
        L.usg_overset_destroy.argtypes = [vp]
        L.usg_overset_iblank.argtypes = [vp, C.c_int, _ip, _ip]
        L.usg_overset_receptors.argtypes = [vp, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _dp]
        L.usg_overset_resolution.argtypes = [vp, C.c_int, _dp, _dp]
        L.usg_overset_median.argtypes = [_ip, _dp, C.c_int]
        L.usg_stencil_apply_host.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_long, C.c_long, C.c_int, _ip, _ip, _ip,
                                             _dp]
        L.usg_stencil_apply_host_noalloc.argtypes = L.usg_stencil_apply_host.argtypes
        L.usg_overset_cell_stencil.argtypes = [vp, C.POINTER(_MeshDesc), C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _dp]
        L.usg_set_overset.argtypes = [vp, _ip, C.c_int, _ip, _ip, _ip, _dp]
        L.usg_dual_time.argtypes = [vp, C.c_double, C.c_int]
        L.usg_dual_time_advance.argtypes = [vp]
        L.usg_set_cfl.argtypes = [vp, C.c_double]
        L.usg_use_fused.argtypes = [vp, C.c_int]
        L.usg_get_fused.argtypes = [vp, _ip]
        L.usg_set_order.argtypes = [vp, C.c_int]
        L.usg_overset_obb.argtypes = [_dp, C.c_int, _dp, _dp, _dp]
        L.usg_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise UsgError(f"usg error {rc}: {lib().usg_last_error().decode()}")


def _desc(m: _mesh.Mesh):
    keep = [np.ascontiguousarray(m.xy, np.float64), np.ascontiguousarray(m.fnode, np.int32),
            np.ascontiguousarray(m.flc, np.int32), np.ascontiguousarray(m.frc, np.int32)]
    d = _MeshDesc(m.nn, m.nf, m.nc, keep[0].ctypes.data_as(_dp), keep[1].ctypes.data_as(_ip),
                  keep[2].ctypes.data_as(_ip), keep[3].ctypes.data_as(_ip))
    return d, keep


def mesh_geometry(m: _mesh.Mesh) -> dict:
    """Host geometry and wave counts of the product (no GPU needed)."""
    d, keep = _desc(m)
    out = {"vol": np.zeros(m.nc), "cc": np.zeros((m.nc, 2)), "fn": np.zeros((m.nf, 2)),
           "fa": np.zeros(m.nf), "fc": np.zeros((m.nf, 2)), "wdist": np.zeros(m.nc)}
    a, b = C.c_int(), C.c_int()
    _check(lib().usg_mesh_build(C.byref(d), *(out[k].ctypes.data_as(_dp) for k in
                                              ("vol", "cc", "fn", "fa", "fc", "wdist")), C.byref(a), C.byref(b)))
    out["nfwd"], out["nbwd"] = a.value, b.value
    return out


def mesh_renumber(m: _mesh.Mesh) -> tuple[_mesh.Mesh, np.ndarray, int]:
    """Colour-blocked renumbering by the library (usg_mesh_renumber):
    (renumbered mesh, perm[nc] new -> old cell, number of colours)."""
    d, keep = _desc(m)
    perm = np.zeros(m.nc, np.int32)
    fnode = np.zeros((m.nf, 2), np.int32)
    flc = np.zeros(m.nf, np.int32)
    frc = np.zeros(m.nf, np.int32)
    ncol = C.c_int()
    _check(lib().usg_mesh_renumber(C.byref(d), perm.ctypes.data_as(_ip), fnode.ctypes.data_as(_ip),
                                   flc.ctypes.data_as(_ip), frc.ctypes.data_as(_ip), C.byref(ncol)))
    return _mesh.Mesh(m.xy.copy(), fnode, flc, frc, m.nc, m.name), perm, ncol.value


def cell_nodes(m: _mesh.Mesh) -> tuple[np.ndarray, np.ndarray]:
    """Cell vertices in the reference's order: (cnode[nc, 4] -1 padded, cnv[nc])."""
    d, keep = _desc(m)
    cnode = np.zeros((m.nc, 4), np.int32)
    cnv = np.zeros(m.nc, np.int32)
    _check(lib().usg_mesh_cells(C.byref(d), cnode.ctypes.data_as(_ip), cnv.ctypes.data_as(_ip)))
    return cnode, cnv


def freestream_state(mach: float, aoa: float) -> np.ndarray:
    """Non-dimensional free stream (rho, p, u, v) of a context at Mach `mach`, `aoa` degrees."""
    fs = np.zeros(4)
    _check(lib().usg_freestream_state(float(mach), float(aoa), fs.ctypes.data_as(_dp)))
    return fs


def _prim_fs(m, U, fs):
    U = np.ascontiguousarray(U, np.float64)
    if U.shape != (m.nc, 4):
        raise UsgError(f"U must be ({m.nc}, 4), got {U.shape}")
    return U, np.ascontiguousarray(fs, np.float64).reshape(4)


def write_cell_data(filename: str, m: _mesh.Mesh, U, fs, proc: int = 0, mesh_tag: int = 0) -> str:
    """writeCellData: Tecplot file "<filename>_<mesh_tag+1>_<proc+1>"; returns its path."""
    U, fs = _prim_fs(m, U, fs)
    d, keep = _desc(m)
    _check(lib().usg_write_cell_data(C.byref(d), U.ctypes.data_as(_dp), fs.ctypes.data_as(_dp),
                                     os.fsencode(filename), proc, mesh_tag))
    return f"{filename}_{mesh_tag + 1}_{proc + 1}"


def surface_cp(m: _mesh.Mesh, U, fs) -> tuple[np.ndarray, np.ndarray]:
    """outPutCp: (table[nwall, 4], curve[ncurve, 4]) of (x, y, signed count, Cp)."""
    U, fs = _prim_fs(m, U, fs)
    d, keep = _desc(m)
    nw, ncv = C.c_int(), C.c_int()
    L = lib()
    _check(L.usg_surface_cp(C.byref(d), U.ctypes.data_as(_dp), fs.ctypes.data_as(_dp), 0, None, None,
                            C.byref(nw), C.byref(ncv)))
    table = np.zeros((nw.value, 4))
    curve = np.zeros((nw.value, 4))
    _check(L.usg_surface_cp(C.byref(d), U.ctypes.data_as(_dp), fs.ctypes.data_as(_dp), nw.value,
                            table.ctypes.data_as(_dp), curve.ctypes.data_as(_dp), C.byref(nw), C.byref(ncv)))
    return table, curve[:ncv.value].copy()


def write_cp(path: str, curve) -> None:
    """The reference's Cp_<step> text file of a curve from surface_cp."""
    c = np.ascontiguousarray(curve, np.float64).reshape(-1, 4)
    _check(lib().usg_write_cp(os.fsencode(path), c.ctypes.data_as(_dp), c.shape[0]))


def overset_connect(meshes, overset_faces=None, mexclude: int = 3, nfringe: int = 1) -> list[dict]:
    """UniTioga connectivity of the blocks (usg_overset_*): per block, node
    and cell iblank, resolutions, and the receptors whose donor cell lies in
    it (recv_block, recv_node, donor_cell, nw, donor_node[n,4], w[n,4]).
    mexclude / nfringe default to MeshBlock's 3 / 1; the reference's input
    file uses 1 / 2."""
    k = len(meshes)
    descs = (_MeshDesc * k)()
    keep = []
    flags = (_ip * k)()
    for i, m in enumerate(meshes):
        d, kp = _desc(m)
        descs[i] = d
        keep.append(kp)
        f = None if overset_faces is None or overset_faces[i] is None else \
            np.ascontiguousarray(overset_faces[i], np.int32)
        keep.append(f)
        flags[i] = f.ctypes.data_as(_ip) if f is not None else _ip()
    L = lib()
    h = C.c_void_p()
    _check(L.usg_overset_create(descs, flags, k, int(mexclude), int(nfringe), C.byref(h)))
    out = []
    try:
        for i, m in enumerate(meshes):
            nib = np.zeros(m.nn, np.int32)
            cib = np.zeros(m.nc, np.int32)
            _check(L.usg_overset_iblank(h, i, nib.ctypes.data_as(_ip), cib.ctypes.data_as(_ip)))
            nres, cres = np.zeros(m.nn), np.zeros(m.nc)
            _check(L.usg_overset_resolution(h, i, nres.ctypes.data_as(_dp), cres.ctypes.data_as(_dp)))
            n = C.c_int()
            _check(L.usg_overset_receptors(h, i, C.byref(n), None, None, None, None, None, None))
            r = {"recv_block": np.zeros(n.value, np.int32), "recv_node": np.zeros(n.value, np.int32),
                 "donor_cell": np.zeros(n.value, np.int32), "nw": np.zeros(n.value, np.int32),
                 "donor_node": np.zeros((n.value, 4), np.int32), "w": np.zeros((n.value, 4))}
            _check(L.usg_overset_receptors(h, i, C.byref(n), *(r[q].ctypes.data_as(_ip) for q in
                                                               ("recv_block", "recv_node", "donor_cell", "nw",
                                                                "donor_node")), r["w"].ctypes.data_as(_dp)))
            out.append({"node_iblank": nib, "cell_iblank": cib, "node_res": nres, "cell_res": cres, **r})
    finally:
        L.usg_overset_destroy(h)
    return out


def overset_coupling(meshes, overset_faces=None, mexclude: int = 3, nfringe: int = 1) -> tuple[list, dict]:
    """Connectivity of the blocks (as overset_connect) and the cell-centred
    coupling stencil of usg_overset_cell_stencil over their cells in block
    order: {"blank", "recv", "off", "donor", "w"}."""
    k = len(meshes)
    descs = (_MeshDesc * k)()
    keep = []
    flags = (_ip * k)()
    for i, m in enumerate(meshes):
        d, kp = _desc(m)
        descs[i] = d
        keep.append(kp)
        f = None if overset_faces is None or overset_faces[i] is None else \
            np.ascontiguousarray(overset_faces[i], np.int32)
        keep.append(f)
        flags[i] = f.ctypes.data_as(_ip) if f is not None else _ip()
    L = lib()
    h = C.c_void_p()
    _check(L.usg_overset_create(descs, flags, k, int(mexclude), int(nfringe), C.byref(h)))
    try:
        conn = []
        for i, m in enumerate(meshes):
            nib = np.zeros(m.nn, np.int32)
            cib = np.zeros(m.nc, np.int32)
            _check(L.usg_overset_iblank(h, i, nib.ctypes.data_as(_ip), cib.ctypes.data_as(_ip)))
            n = C.c_int()
            _check(L.usg_overset_receptors(h, i, C.byref(n), None, None, None, None, None, None))
            r = {"recv_block": np.zeros(n.value, np.int32), "recv_node": np.zeros(n.value, np.int32),
                 "donor_cell": np.zeros(n.value, np.int32), "nw": np.zeros(n.value, np.int32),
                 "donor_node": np.zeros((n.value, 4), np.int32), "w": np.zeros((n.value, 4))}
            _check(L.usg_overset_receptors(h, i, C.byref(n), *(r[q].ctypes.data_as(_ip) for q in
                                                               ("recv_block", "recv_node", "donor_cell", "nw",
                                                                "donor_node")), r["w"].ctypes.data_as(_dp)))
            conn.append({"node_iblank": nib, "cell_iblank": cib, **r})
        ntot = sum(m.nc for m in meshes)
        nr, nd = C.c_int(), C.c_int()
        blank = np.zeros(ntot, np.int32)
        _check(L.usg_overset_cell_stencil(h, descs, k, blank.ctypes.data_as(_ip), C.byref(nr), None, None,
                                          C.byref(nd), None, None))
        st = {"blank": blank, "recv": np.zeros(nr.value, np.int32), "off": np.zeros(nr.value + 1, np.int32),
              "donor": np.zeros(nd.value, np.int32), "w": np.zeros(nd.value)}
        _check(L.usg_overset_cell_stencil(h, descs, k, None, C.byref(nr), st["recv"].ctypes.data_as(_ip),
                                          st["off"].ctypes.data_as(_ip), C.byref(nd),
                                          st["donor"].ctypes.data_as(_ip), st["w"].ctypes.data_as(_dp)))
    finally:
        L.usg_overset_destroy(h)
    return conn, st


def overset_median(x) -> tuple[np.ndarray, np.ndarray]:
    """find_median_wrap on a copy of x: (partitioned x, companion indices)."""
    x = np.array(x, np.float64)
    ix = np.arange(x.size, dtype=np.int32)
    _check(lib().usg_overset_median(ix.ctypes.data_as(_ip), x.ctypes.data_as(_dp), x.size))
    return x, ix


def overset_obb(pts) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """findOBB of points pts[n, 2]: (centre, half sizes, axes as rows)."""
    p = np.ascontiguousarray(pts, np.float64).reshape(-1, 2)
    xc, dxc, vec = np.zeros(2), np.zeros(2), np.zeros((2, 2))
    _check(lib().usg_overset_obb(p.ctypes.data_as(_dp), p.shape[0], xc.ctypes.data_as(_dp),
                                 dxc.ctypes.data_as(_dp), vec.ctypes.data_as(_dp)))
    return xc, dxc, vec


ROW, COLUMN = 0, 1


def nodal_weights(nvert, xv, xp) -> np.ndarray:
    """Donor weights w[n, 4] of points xp[n, 2] in cells xv[n, 4, 2] of
    nvert[n] (3 or 4) vertices (computeNodalWeight2D)."""
    nvert = np.ascontiguousarray(nvert, np.int32)
    n = nvert.size
    X = np.zeros((n, 4, 2))
    X[:, :np.asarray(xv).shape[1]] = xv
    X = np.ascontiguousarray(X)
    P = np.ascontiguousarray(xp, np.float64).reshape(n, 2)
    w = np.zeros((n, 4))
    _check(lib().usg_nodal_weights(n, nvert.ctypes.data_as(_ip), X.ctypes.data_as(_dp), P.ctypes.data_as(_dp),
                                   w.ctypes.data_as(_dp)))
    return w


def stencil_apply(src_ptr, dst_ptr, nvar, layout, src_pts, dst_pts, nrecv, recv_ptr, nw_ptr, donor_ptr, w_ptr,
                  stream=None):
    """usg_stencil_apply on device pointers (ints)."""
    _check(lib().usg_stencil_apply(src_ptr, dst_ptr, nvar, layout, src_pts, dst_pts, nrecv, recv_ptr, nw_ptr,
                                   donor_ptr, w_ptr, stream))


class Solver:
    """One mesh + flow strategy + integrator on the GPU (WrapSolver equivalent)."""

    def __init__(self, m: _mesh.Mesh, flow: int = EULER, mach: float = 0.5, aoa: float = 0.0,
                 reynolds: float = 1e6, precision: int = FP64, cfl_lusgs: float = 400.0, cfl_rk: float = 0.8):
        self.mesh = m
        self.nc = m.nc
        self.flow = flow
        self.ncase = 1
        self.case_cell = np.array([0, m.nc])
        d, keep = _desc(m)
        p = _Params(flow, precision, mach, aoa, reynolds, cfl_lusgs, cfl_rk)
        h = C.c_void_p()
        _check(lib().usg_ctx_create(C.byref(d), C.byref(p), C.byref(h)))
        self._h = h
        self.meshes = [m]
        self.case_params = [dict(mach=float(mach), aoa=float(aoa), reynolds=float(reynolds))]
        self._common(flow, precision, cfl_lusgs, cfl_rk)

    def _common(self, flow, precision, cfl_lusgs, cfl_rk):
        self.precision = precision
        self.cfl = (float(cfl_lusgs), float(cfl_rk))
        self.step = 0                    # pseudo-time iterations taken
        self.hist_step = np.zeros(0, np.int64)          # iterations with a recorded residual
        self.hist = np.zeros((0, self.ncase))           # their residual norms per case

    def close(self):
        if getattr(self, "_h", None):
            lib().usg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- state -----------------------------------------------------------
    def init_freestream(self):
        _check(lib().usg_init_freestream(self._h))

    def set_prim(self, U, turb=None):
        U = np.ascontiguousarray(U, np.float64).reshape(self.nc, 4)
        tp = None
        if turb is not None:
            turb = np.ascontiguousarray(turb, np.float64).reshape(self.nc, 2)
            tp = turb.ctypes.data_as(_dp)
        _check(lib().usg_set_prim(self._h, U.ctypes.data_as(_dp), tp))

    def set_state(self, U, W=None, turb=None, tcons=None, mut=None):
        """usg_set_state: restart from a full state (W, tcons may be None)."""
        arrs = []

        def p(a, w):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float64).reshape(self.nc, w)
            arrs.append(a)
            return a.ctypes.data_as(_dp)
        _check(lib().usg_set_state(self._h, p(U, 4), p(W, 4), p(turb, 2), p(tcons, 2), p(mut, 1)))

    def field(self, fid: int) -> np.ndarray:
        if fid == F_UEDGE:
            nf = sum(m.nf for m in self.meshes)
            out = np.zeros(nf * 4)
            _check(lib().usg_get_field(self._h, fid, out.ctypes.data_as(_dp), out.size))
            return out.reshape(nf, 4)
        w = _FIELD_WIDTH[fid]
        out = np.zeros(self.nc * w)
        _check(lib().usg_get_field(self._h, fid, out.ctypes.data_as(_dp), out.size))
        return out.reshape(self.nc, w) if w > 1 else out

    @property
    def U(self):
        return self.field(F_PRIM)

    @property
    def R(self):
        return self.field(F_RES)

    # --- work ------------------------------------------------------------
    def residual(self):
        _check(lib().usg_residual(self._h))

    def _steps(self, fn, n, history):
        h = np.zeros(max(n, 1) * self.ncase) if history else None
        _check(fn(self._h, n, h.ctypes.data_as(_dp) if history else None))
        first = self.step + 1
        self.step += n
        if not history:
            return None
        h = h[:n * self.ncase].reshape(n, self.ncase)
        self.hist_step = np.concatenate([self.hist_step, np.arange(first, first + n)])
        self.hist = np.concatenate([self.hist, h])
        return h if isinstance(self, BatchSolver) else h[:, 0]

    def lusgs(self, n: int = 1, history: bool = False):
        """n LU-SGS iterations; with history, ||R_rho/V||_2 per iteration."""
        return self._steps(lib().usg_lusgs_step, n, history)

    def rk5(self, n: int = 1, history: bool = False):
        return self._steps(lib().usg_rk5_step, n, history)

    def res_l2(self):
        v = np.zeros(self.ncase)
        _check(lib().usg_res_l2(self._h, v.ctypes.data_as(_dp)))
        return v if isinstance(self, BatchSolver) else float(v[0])

    def forces(self):
        cl, cd = np.zeros(self.ncase), np.zeros(self.ncase)
        _check(lib().usg_forces(self._h, cl.ctypes.data_as(_dp), cd.ctypes.data_as(_dp)))
        if isinstance(self, BatchSolver):
            return cl, cd
        return float(cl[0]), float(cd[0])

    def dw_rho_sq(self) -> np.ndarray:
        """The reference's LU-SGS monitor (LUSGSIntegrator::postprocessUpdate,
        LUSGS_Integrator.cpp:131-153): sum over a case's cells of the last
        iteration's dW_rho squared, added in cell order; one value per case."""
        d = self.field(F_DW)[:, 0]
        c = self.case_cell
        return np.array([float(np.cumsum(d[c[i]:c[i + 1]] * d[c[i]:c[i + 1]])[-1]) if c[i + 1] > c[i] else 0.0
                         for i in range(self.ncase)])

    def rk_monitor(self) -> np.ndarray:
        """The reference's RK5 monitor (RungeKuttaStrategy::converged,
        RungeKuttaStrategy.h:87-105): sum over a case's cells of
        |rho - rho at the start of the last step|, added in cell order."""
        d = np.abs(self.U[:, 0] - self.field(F_W0)[:, 0])
        c = self.case_cell
        return np.array([float(np.cumsum(d[c[i]:c[i + 1]])[-1]) if c[i + 1] > c[i] else 0.0
                         for i in range(self.ncase)])

    def waves(self):
        a, b = C.c_int(), C.c_int()
        _check(lib().usg_sweep_waves(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def sync(self):
        _check(lib().usg_sync(self._h))

    def graphs(self, on: bool):
        _check(lib().usg_use_graphs(self._h, int(on)))

    def fused(self, on: bool):
        """Face fluxes and cell gathers in one launch (FP64; default off)."""
        _check(lib().usg_use_fused(self._h, int(on)))

    def is_fused(self) -> bool:
        v = C.c_int()
        _check(lib().usg_get_fused(self._h, C.byref(v)))
        return bool(v.value)

    def timing(self, on: bool):
        _check(lib().usg_timing_enable(self._h, int(on)))

    def kernel_time(self):
        t, n = C.c_double(), C.c_long()
        _check(lib().usg_kernel_time(self._h, C.byref(t), C.byref(n)))
        return t.value, n.value

    # --- outputs and checkpoints ----------------------------------------
    def freestream(self, case: int = 0) -> np.ndarray:
        p = self.case_params[case]
        return freestream_state(p["mach"], p["aoa"])

    def case_U(self, case: int = 0) -> np.ndarray:
        c = self.case_cell
        return self.U[int(c[case]):int(c[case + 1])]

    def write_cell_data(self, filename: str, proc: int = 0, mesh_tag: int = 0, case: int = 0) -> str:
        """writeCellData of one case's current state; returns the file path."""
        return write_cell_data(filename, self.meshes[case], self.case_U(case), self.freestream(case), proc,
                               mesh_tag)

    def surface_cp(self, case: int = 0):
        """outPutCp of one case's current state: (table, curve), see surface_cp."""
        return surface_cp(self.meshes[case], self.case_U(case), self.freestream(case))

    def write_cp(self, path: str, case: int = 0) -> np.ndarray:
        _, curve = self.surface_cp(case)
        write_cp(path, curve)
        return curve

    def save(self, path: str) -> None:
        """Checkpoint: meshes, parameters, the pseudo-time state (U, W, and
        for SST k, omega, rho k, rho omega, mu_t), iteration count and
        residual history, in one .npz file."""
        d = {"flow": self.flow, "precision": self.precision, "cfl": np.array(self.cfl), "step": self.step,
             "ncase": self.ncase, "batch": isinstance(self, BatchSolver),
             "mach": np.array([p["mach"] for p in self.case_params]),
             "aoa": np.array([p["aoa"] for p in self.case_params]),
             "reynolds": np.array([p["reynolds"] for p in self.case_params]),
             "U": self.U, "W": self.field(F_CONS), "hist_step": self.hist_step, "hist": self.hist}
        if self.flow == SST:
            d.update(turb=self.field(F_TURB), tcons=self.field(F_TCONS), mut=self.field(F_MUT))
        for i, m in enumerate(self.meshes):
            d.update({f"xy{i}": m.xy, f"fnode{i}": m.fnode, f"flc{i}": m.flc, f"frc{i}": m.frc,
                      f"nc{i}": m.nc, f"name{i}": m.name})
        with open(path, "wb") as f:
            np.savez(f, **d)


    # --- overset coupling and dual time ---------------------------------
    def set_overset(self, st: dict | None) -> None:
        """usg_set_overset with a stencil of overset_coupling (None removes it)."""
        L = lib()
        if st is None:
            _check(L.usg_set_overset(self._h, None, 0, None, None, None, None))
            return
        arrs = [np.ascontiguousarray(st[q], np.int32) for q in ("blank", "recv", "off", "donor")]
        w = np.ascontiguousarray(st["w"], np.float64)
        _check(L.usg_set_overset(self._h, arrs[0].ctypes.data_as(_ip), arrs[1].size, arrs[1].ctypes.data_as(_ip),
                                 arrs[2].ctypes.data_as(_ip), arrs[3].ctypes.data_as(_ip), w.ctypes.data_as(_dp)))

    def set_cfl(self, cfl: float) -> None:
        """LU-SGS pseudo-time CFL number from the next iteration on."""
        _check(lib().usg_set_cfl(self._h, float(cfl)))
        self.cfl = (float(cfl), self.cfl[1])

    def set_order(self, order: int) -> None:
        """Reconstruction order from the next residual on (2: limited linear, 1: cell values)."""
        _check(lib().usg_set_order(self._h, int(order)))

    def dual_time(self, dt_phys: float, order: int = 2) -> None:
        _check(lib().usg_dual_time(self._h, float(dt_phys), int(order)))

    def dual_time_advance(self) -> None:
        _check(lib().usg_dual_time_advance(self._h))

    # --- launch configuration -------------------------------------------
    def blocks(self) -> list:
        out = []
        for g in range(4):
            v = C.c_int()
            _check(lib().usg_get_block(self._h, g, C.byref(v)))
            out.append(v.value)
        return out

    def set_blocks(self, blocks) -> None:
        for g, t in enumerate(blocks):
            _check(lib().usg_set_block(self._h, g, int(t)))

    def launch_config(self) -> list:
        """[threads per block of the gradient, face, cell and sweep groups,
        1 if the fused face-cell launch is in use else 0]."""
        return self.blocks() + [int(self.is_fused())]

    def set_launch_config(self, cfg) -> None:
        self.set_blocks(cfg[:4])
        if len(cfg) > 4:
            self.fused(bool(cfg[4]))

    def autotune(self, iters: int = 5) -> list:
        """Runs usg_autotune (block sizes and, in FP64, the launch scheme);
        returns the chosen launch_config()."""
        _check(lib().usg_autotune(self._h, iters, None))
        return self.launch_config()


def load_checkpoint(path: str, precision: int | None = None) -> "Solver":
    """A Solver (or BatchSolver) restored from Solver.save: same meshes and
    parameters, state loaded with usg_set_prim, iteration count and history
    kept.  Continuing it gives the same iterates as the run that saved it."""
    with np.load(path, allow_pickle=False) as z:
        k = int(z["ncase"])
        meshes = [_mesh.Mesh(z[f"xy{i}"], z[f"fnode{i}"], z[f"flc{i}"], z[f"frc{i}"], int(z[f"nc{i}"]),
                             str(z[f"name{i}"])) for i in range(k)]
        flow = int(z["flow"])
        prec = int(z["precision"]) if precision is None else precision
        cfl = z["cfl"]
        if bool(z["batch"]):
            s = BatchSolver(meshes, flow, z["mach"], z["aoa"], z["reynolds"], prec, cfl[0], cfl[1])
        else:
            s = Solver(meshes[0], flow, float(z["mach"][0]), float(z["aoa"][0]), float(z["reynolds"][0]), prec,
                       cfl[0], cfl[1])
        sst = flow == SST
        s.set_state(z["U"], z["W"], z["turb"] if sst else None, z["tcons"] if sst else None,
                    z["mut"] if sst else None)
        s.step = int(z["step"])
        s.hist_step = z["hist_step"].copy()
        s.hist = z["hist"].copy()
    return s


class BatchSolver(Solver):
    """Many independent cases in one context (usg_batch_create): per-case
    free streams, one launch sequence per iteration for all of them.  Fields
    are case-major over the union; per-case outputs are arrays."""

    def __init__(self, meshes, flow: int = EULER, mach=0.5, aoa=0.0, reynolds=1e6, precision: int = FP64,
                 cfl_lusgs: float = 400.0, cfl_rk: float = 0.8):
        k = len(meshes)
        mach, aoa, reynolds = (np.broadcast_to(np.asarray(v, np.float64), (k,)) for v in (mach, aoa, reynolds))
        descs = (_MeshDesc * k)()
        self._keep = []
        for i, m in enumerate(meshes):
            d, keep = _desc(m)
            descs[i] = d
            self._keep.append(keep)
        params = (_Params * k)(*[_Params(flow, precision, float(mach[i]), float(aoa[i]), float(reynolds[i]),
                                         cfl_lusgs, cfl_rk) for i in range(k)])
        h = C.c_void_p()
        _check(lib().usg_batch_create(descs, params, k, C.byref(h)))
        self._h = h
        self.meshes = list(meshes)
        self.flow = flow
        n = C.c_int()
        off = np.zeros(k + 1, np.int32)
        _check(lib().usg_case_count(h, C.byref(n), off.ctypes.data_as(_ip)))
        self.ncase = n.value
        self.case_cell = off.astype(np.int64)
        self.nc = int(off[-1])
        self.case_params = [dict(mach=float(mach[i]), aoa=float(aoa[i]), reynolds=float(reynolds[i]))
                            for i in range(k)]
        self._common(flow, precision, cfl_lusgs, cfl_rk)

    def status(self) -> np.ndarray:
        """Per-case status: 0 running, bit 0 HLLC p* check failed, bit 1 slip-wall check failed."""
        st = np.zeros(self.ncase, np.int32)
        _check(lib().usg_case_status(self._h, st.ctypes.data_as(_ip)))
        return st

    def case_slice(self, i: int) -> slice:
        return slice(int(self.case_cell[i]), int(self.case_cell[i + 1]))
