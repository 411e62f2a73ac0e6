"""ctypes front end of the CPU FP64 oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry and the
cpu_baseline leg of bench.py, never by the product package.  The arrays it
exposes are views of the oracle's own AoS buffers, laid out as in the
reference (``U[cell*4 + var]``, ``gradU[(cell*4 + var)*2 + dim]``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")

EULER, LAMINAR, SST = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _Mesh(C.Structure):
    _fields_ = [("nn", C.c_int), ("nf", C.c_int), ("nc", C.c_int),
                ("xy", _dp), ("fnode", _ip), ("flc", _ip), ("frc", _ip),
                ("cnode", _ip), ("cnv", _ip), ("cface", _ip), ("cnf", _ip),
                ("vol", _dp), ("cc", _dp), ("fn", _dp), ("fa", _dp), ("fc", _dp),
                ("wdist", _dp), ("wdist_wall", _dp)]


_SOLVER_DOUBLES_A = ["R_gas", "Gamma", "PI", "fs_mach", "fs_AOA", "fs_density", "fs_pressure",
                     "fs_eigen_Length", "fs_soundSpeed", "fs_velocity_magnitude"]


class _Solver(C.Structure):
    _fields_ = ([("m", C.POINTER(_Mesh)), ("flow", C.c_int), ("neq", C.c_int)]
                + [(n, C.c_double) for n in _SOLVER_DOUBLES_A]
                + [("fs_velocity_components", C.c_double * 3)]
                + [(n, C.c_double) for n in ["fs_Temperature", "fs_Pr", "fs_Prt", "fs_mu", "fs_Re",
                                             "fsnd_density", "fsnd_pressure", "fsnd_soundSpeed"]]
                + [("fsnd_velocity_components", C.c_double * 3), ("fs_primVar", C.c_double * 5)]
                + [(n, _dp) for n in ["U", "U_edge", "gradU", "gradT", "R", "lamc", "lamv", "dt",
                                      "Umax", "Umin", "phi", "UL", "UR", "mu", "mu_edge"]]
                + [("cfl_lusgs", C.c_double)]
                + [(n, _dp) for n in ["D", "W", "dW1", "dW"]]
                + [("cfl_rk", C.c_double), ("rk_alpha", C.c_double * 5)]
                + [(n, _dp) for n in ["W0", "Wnow", "tU", "tUedge", "tgrad", "F1", "F2", "mut",
                                      "mut_face", "tFc", "tFv", "tR", "tD", "tW", "tdW1", "tdW"]]
                + [(n, C.c_double) for n in ["fs_mut", "fsnd_mut", "fsnd_mu"]]
                + [("fs_turb", C.c_double * 2), ("fsnd_turb", C.c_double * 2)]
                + [("hllc_pstar_violations", C.c_long), ("Fv_face", _dp), ("Fv_all", _dp)]
                + [("dt_phys", C.c_double), ("dt_order", C.c_int), ("dt_steps", C.c_int),
                   ("dts", C.c_double * 3), ("Wt1", _dp), ("Wt2", _dp), ("blank", _ip), ("order", C.c_int)])


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < max(
            os.path.getmtime(os.path.join(HERE, f)) for f in ("usg_oracle.c", "usg_oracle.h")):
        subprocess.run(["make", "-B", "-C", HERE, "liborc.so"], check=True, capture_output=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_nodal_weights.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_stencil_apply.argtypes = [_dp, _dp, C.c_int, C.c_int, _ip, _ip, _ip, _dp]
        L.orc_mesh_create.restype = C.POINTER(_Mesh)
        L.orc_mesh_create.argtypes = [C.c_int, _dp, C.c_int, _ip, _ip, _ip, C.c_int]
        L.orc_mesh_destroy.argtypes = [C.POINTER(_Mesh)]
        L.orc_solver_create.restype = C.POINTER(_Solver)
        L.orc_solver_create.argtypes = [C.POINTER(_Mesh), C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_solver_destroy.argtypes = [C.POINTER(_Solver)]
        for fn in ("orc_init_data", "orc_residual", "orc_lusgs_step", "orc_rk5_step"):
            getattr(L, fn).argtypes = [C.POINTER(_Solver)]
        L.orc_set_prim.argtypes = [C.POINTER(_Solver), _dp]
        L.orc_density_residual_l2.restype = C.c_double
        L.orc_density_residual_l2.argtypes = [C.POINTER(_Solver)]
        L.orc_dw_rho_sq.restype = C.c_double
        L.orc_dw_rho_sq.argtypes = [C.POINTER(_Solver)]
        L.orc_forces.argtypes = [C.POINTER(_Solver), _dp, _dp]
        L.orc_sst_eddy_viscosity.argtypes = [C.POINTER(_Solver)]
        L.orc_dual_time.argtypes = [C.POINTER(_Solver), C.c_double, C.c_int]
        L.orc_dual_time_advance.argtypes = [C.POINTER(_Solver)]
        L.orc_set_blank.argtypes = [C.POINTER(_Solver), _ip]
        L.orc_set_cells.argtypes = [C.POINTER(_Solver), C.c_int, _ip, _dp]
        L.orc_set_order.argtypes = [C.POINTER(_Solver), C.c_int]
        L.orc_hllc_face.argtypes = [_dp, _dp, _dp, C.c_double, _dp, _ip]
        L.orc_lusgs_sweeps.argtypes = [C.POINTER(_Solver)]
        L.orc_cell_volume.restype = C.c_double
        L.orc_cell_volume.argtypes = [_dp, C.c_int]
        L.orc_prim_to_cons.argtypes = [_dp, _dp]
        L.orc_cons_to_prim.argtypes = [_dp, _dp]
        L.orc_lusgs_levels.argtypes = [C.POINTER(_Mesh), _ip, _ip, _ip, _ip]
        _lib = L
    return _lib


def _arr(ptr, n, dtype=np.float64):
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype) if n else np.zeros(0, dtype)


class OracleMesh:
    """Geometry of a mesh as the reference computes it."""

    def __init__(self, mesh):
        L = lib()
        self._keep = [np.ascontiguousarray(mesh.xy, dtype=np.float64),
                      np.ascontiguousarray(mesh.fnode, dtype=np.int32),
                      np.ascontiguousarray(mesh.flc, dtype=np.int32),
                      np.ascontiguousarray(mesh.frc, dtype=np.int32)]
        xy, fnode, flc, frc = self._keep
        self.ptr = L.orc_mesh_create(mesh.nn, xy.ctypes.data_as(_dp), mesh.nf, fnode.ctypes.data_as(_ip),
                                     flc.ctypes.data_as(_ip), frc.ctypes.data_as(_ip), mesh.nc)
        m = self.ptr.contents
        self.nc, self.nf = m.nc, m.nf
        self.vol = _arr(m.vol, m.nc)
        self.cc = _arr(m.cc, 2 * m.nc).reshape(m.nc, 2)
        self.fn = _arr(m.fn, 2 * m.nf).reshape(m.nf, 2)
        self.fa = _arr(m.fa, m.nf)
        self.fc = _arr(m.fc, 2 * m.nf).reshape(m.nf, 2)
        self.wdist = _arr(m.wdist, m.nc)
        self.wdist_wall = _arr(m.wdist_wall, m.nc)
        self.cnode = np.ctypeslib.as_array(m.cnode, shape=(4 * m.nc,)).reshape(m.nc, 4)
        self.cnv = np.ctypeslib.as_array(m.cnv, shape=(m.nc,))
        self.cface = np.ctypeslib.as_array(m.cface, shape=(4 * m.nc,)).reshape(m.nc, 4)
        self.cnf = np.ctypeslib.as_array(m.cnf, shape=(m.nc,))

    def levels(self):
        fl = np.zeros(self.nc, np.int32)
        bl = np.zeros(self.nc, np.int32)
        a, b = C.c_int(), C.c_int()
        lib().orc_lusgs_levels(self.ptr, fl.ctypes.data_as(_ip), bl.ctypes.data_as(_ip), C.byref(a), C.byref(b))
        return fl, bl, a.value, b.value

    def __del__(self):
        try:
            lib().orc_mesh_destroy(self.ptr)
        except Exception:
            pass


class OracleSolver:
    """The reference's serial solver (Euler2D / ViscousFlow2D with LU-SGS or RK5)."""

    def __init__(self, omesh: OracleMesh, flow: int = EULER, mach: float = 0.5, aoa: float = 0.0,
                 re: float = 1e6):
        self.mesh = omesh
        self.ptr = lib().orc_solver_create(omesh.ptr, flow, mach, aoa, re)
        self.s = self.ptr.contents
        self.flow = flow
        lib().orc_init_data(self.ptr)

    def field(self, name: str, per: int, n: str = "nc"):
        cnt = getattr(self.mesh, n)
        return _arr(getattr(self.s, name), per * cnt).reshape(cnt, per) if per > 1 else _arr(getattr(self.s, name), cnt)

    @property
    def U(self):
        return self.field("U", 4)

    @property
    def R(self):
        return self.field("R", 4)

    def init_data(self):
        """Free-stream cold start (orc_init_data)."""
        lib().orc_init_data(self.ptr)

    def set_prim(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        lib().orc_set_prim(self.ptr, U.ctypes.data_as(_dp))

    def residual(self):
        lib().orc_residual(self.ptr)

    def lusgs_step(self):
        lib().orc_lusgs_step(self.ptr)

    def rk5_step(self):
        lib().orc_rk5_step(self.ptr)

    def sweeps(self):
        lib().orc_lusgs_sweeps(self.ptr)

    def res_l2(self) -> float:
        return lib().orc_density_residual_l2(self.ptr)

    def dual_time(self, dt_phys: float, order: int = 2):
        """Physical time step dt_phys by BDF1 on the first step, then BDF``order``."""
        lib().orc_dual_time(self.ptr, float(dt_phys), int(order))

    def dual_time_advance(self):
        lib().orc_dual_time_advance(self.ptr)

    def set_blank(self, blank):
        """Overset cell blanking (1 field, 0 hole, -1 receptor; None clears)."""
        if blank is None:
            lib().orc_set_blank(self.ptr, None)
            return
        b = np.ascontiguousarray(blank, np.int32)
        lib().orc_set_blank(self.ptr, b.ctypes.data_as(_ip))

    def set_order(self, order: int):
        lib().orc_set_order(self.ptr, int(order))

    def set_cells(self, idx, U):
        """Primitive states U[n, 4] of cells idx (W from U), e.g. overset receptors."""
        idx = np.ascontiguousarray(idx, np.int32)
        U = np.ascontiguousarray(U, np.float64).reshape(idx.size, 4)
        lib().orc_set_cells(self.ptr, idx.size, idx.ctypes.data_as(_ip), U.ctypes.data_as(_dp))

    def eddy_viscosity(self):
        """mu_t of the current state (tW = rho tU, the last residual's F2)."""
        lib().orc_sst_eddy_viscosity(self.ptr)

    def forces(self):
        cl, cd = C.c_double(), C.c_double()
        lib().orc_forces(self.ptr, C.byref(cl), C.byref(cd))
        return cl.value, cd.value

    def __del__(self):
        try:
            lib().orc_solver_destroy(self.ptr)
        except Exception:
            pass


def hllc_face(UL, UR, n, area):
    f = np.zeros(4)
    ok = C.c_int()
    UL = np.ascontiguousarray(UL, np.float64)
    UR = np.ascontiguousarray(UR, np.float64)
    n = np.ascontiguousarray(n, np.float64)
    lib().orc_hllc_face(UL.ctypes.data_as(_dp), UR.ctypes.data_as(_dp), n.ctypes.data_as(_dp), float(area),
                        f.ctypes.data_as(_dp), C.byref(ok))
    return f, bool(ok.value)


def prim_to_cons(U):
    U = np.ascontiguousarray(U, np.float64)
    W = np.zeros(4)
    lib().orc_prim_to_cons(U.ctypes.data_as(_dp), W.ctypes.data_as(_dp))
    return W


def cons_to_prim(W):
    W = np.ascontiguousarray(W, np.float64)
    U = np.zeros(4)
    lib().orc_cons_to_prim(W.ctypes.data_as(_dp), U.ctypes.data_as(_dp))
    return U


def cell_volume(xy):
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    return lib().orc_cell_volume(xy.ctypes.data_as(_dp), xy.shape[0])


def nodal_weights(nvert, xv, xp):
    """orc_nodal_weights of one point: weights[nvert] (or None on a degenerate cell)."""
    xv = np.ascontiguousarray(xv, np.float64).reshape(-1)
    xp = np.ascontiguousarray(xp, np.float64).reshape(2)
    w = np.zeros(4)
    rc = lib().orc_nodal_weights(int(nvert), xv.ctypes.data_as(_dp), xp.ctypes.data_as(_dp), w.ctypes.data_as(_dp))
    return None if rc else w[:nvert]


def stencil_apply(src, dst, recv, nw, donor, w):
    """orc_stencil_apply, ROW layout: src[npts, nvar] -> dst[npts', nvar] in place."""
    src = np.ascontiguousarray(src, np.float64)
    nvar = src.shape[1]
    recv, nw, donor = (np.ascontiguousarray(a, np.int32) for a in (recv, nw, donor))
    w = np.ascontiguousarray(w, np.float64)
    lib().orc_stencil_apply(src.ctypes.data_as(_dp), dst.ctypes.data_as(_dp), nvar, recv.size,
                            recv.ctypes.data_as(_ip), nw.ctypes.data_as(_ip), donor.ctypes.data_as(_ip),
                            w.ctypes.data_as(_dp))
    return dst
