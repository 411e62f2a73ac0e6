"""Pins of the oracle's viscous, turbulence and force functions against
closed forms (no GPU).

Every case is a field for which the reference's discrete operators are
exact, so the textbook formula -- not the oracle's expression -- gives the
expected value: on a mesh of uniform rectangles, Green-Gauss with face
averages reproduces the gradient of a linear field in every cell without a
boundary face, and the face-gradient correction
(LaminarFluxStrategy.h:134-146) vanishes.  The expected values come from

* Stokes' stress tau = mu (grad u + grad u^T - 2/3 div u I) and Fourier's
  heat flux q = -(mu c_p / Pr) grad T with c_p = gamma / (gamma - 1) in the
  non-dimensional variables T = p / rho (LaminarFluxStrategy.h:110-216);
* Sutherland's law with S = 110.5 K (LaminarFluxStrategy.h:76-96);
* Menter's SST model (Menter 1994, AIAA J. 32(8)): the blending functions
  F1/F2, the eddy viscosity, production, destruction and cross diffusion,
  and the wall omega = 60 nu / (beta1 d^2), with the constants of
  SSTKWStrategy.h:13-29 (SSTKWStrategy.h:616-1013);
* the spectral radii of ViscousFlow2D.cpp:220-252 on a uniform state;
* the divergence theorem for the integrated wall force.

Each test names the slip it catches.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tmp_amd import mesh

G = 1.4
PR = 0.72
# SSTKWStrategy.h:13-29
BETA1, BETA2, BSTAR = 0.075, 0.0828, 0.09
SIGW2 = 0.856
CW1, CW2 = 0.533, 0.440
A1 = 0.31
T_INF = 101325.0 / (287.0 * 1.225)   # free-stream temperature (K) of WrapSolver's set_fs_variable2


def sutherland(T, mach, re):
    """Non-dimensional Sutherland viscosity, mu(1) = sqrt(gamma) M / Re."""
    S = 110.5 / T_INF
    return np.sqrt(G) * mach / re * T ** 1.5 * (1 + S) / (T + S)


def interior_masks(m):
    """Cells with no boundary face, and interior faces between two such cells."""
    bcell = np.zeros(m.nc, bool)
    bcell[m.flc[m.frc < 0]] = True
    inner_face = (m.frc >= 0)
    deep_face = inner_face.copy()
    deep_face[inner_face] = ~bcell[m.flc[inner_face]] & ~bcell[m.frc[inner_face]]
    return ~bcell, deep_face


def stokes(ux, uy, vx, vy, mu):
    div = ux + vy
    txx = mu * (2 * ux - 2.0 / 3.0 * div)
    tyy = mu * (2 * vy - 2.0 / 3.0 * div)
    txy = mu * (uy + vx)
    return txx, tyy, txy


# --------------------------------------------------------------------------
# laminar viscous flux (row a13)
# --------------------------------------------------------------------------

def test_stokes_stress_and_fourier_heat_flux():
    """Linear velocity and temperature fields: every deep interior face's
    viscous flux equals (0, (u.tau - q).n A, tau.n A) with Stokes' tau at
    the face viscosity mu(T_face) and Fourier's q.  Catches a wrong 2/3, a
    missing Stokes term, tau_xy from only one shear derivative, swapped
    momentum components, q with the wrong sign or without c_p/Pr, or the
    energy flux without the work term u.tau."""
    mach, re = 0.3, 800.0
    m = mesh.rect(10, 8, lx=1.0, ly=0.8)
    om = O.OracleMesh(m)
    s = O.OracleSolver(om, O.LAMINAR, mach, 0.0, re)
    x, y = om.cc[:, 0], om.cc[:, 1]
    ux, uy, vx, vy, Tx, Ty = 0.5, -0.7, 0.4, 0.9, 0.2, -0.1
    U = np.column_stack([np.ones(m.nc), 1.0 + Tx * x + Ty * y, 0.3 + ux * x + uy * y, -0.2 + vx * x + vy * y])
    s.set_prim(U)
    s.residual()
    Fv = s.field("Fv_all", 4, "nf")
    _, deep = interior_masks(m)
    assert deep.sum() > 50
    checked = 0
    for f in np.nonzero(deep)[0]:
        xf, yf = om.fc[f]
        n, A = om.fn[f], om.fa[f]
        mu = sutherland(1.0 + Tx * xf + Ty * yf, mach, re)
        txx, tyy, txy = stokes(ux, uy, vx, vy, mu)
        k = mu * G / (G - 1) / PR
        qx, qy = -k * Tx, -k * Ty
        u, v = 0.3 + ux * xf + uy * yf, -0.2 + vx * xf + vy * yf
        expect = np.array([0.0,
                           ((u * txx + v * txy - qx) * n[0] + (u * txy + v * tyy - qy) * n[1]) * A,
                           (txx * n[0] + txy * n[1]) * A,
                           (txy * n[0] + tyy * n[1]) * A])
        assert np.allclose(Fv[f], expect, rtol=1e-11, atol=1e-13 * np.abs(expect).max()), (f, Fv[f], expect)
        checked += 1
    assert checked == deep.sum()


def test_sutherland_temperature_dependence():
    """Face viscosity at temperatures away from the free stream follows
    Sutherland's law mu ~ T^1.5 (1 + S)/(T + S), S = 110.5 K / T_inf (a
    dropped (T + S) factor or a T^0.5 fails)."""
    mach, re = 0.3, 800.0
    m = mesh.rect(6, 3)
    om = O.OracleMesh(m)
    s = O.OracleSolver(om, O.LAMINAR, mach, 0.0, re)
    T = 0.6 + 0.8 * om.cc[:, 0]
    U = np.column_stack([np.ones(m.nc) * 1.3, 1.3 * T, np.zeros(m.nc), np.zeros(m.nc)])
    s.set_prim(U)
    s.residual()
    assert np.allclose(s.field("mu", 1), sutherland(T, mach, re), rtol=1e-14, atol=0)


def test_couette_wall_shear_and_adiabatic_wall():
    """u = a y over a no-slip wall at y = 0: the wall faces carry the wall
    shear tau_xy = mu a (momentum flux (-mu a A, 0) along the wall normal
    (0, -1)), no energy flux (u = 0 at the wall and grad T = 0 there,
    LaminarFluxStrategy.h:147-162).  Catches a wall gradient taken from the
    face instead of the cell, or a heat flux at the adiabatic wall."""
    mach, re, a = 0.2, 500.0, 0.8
    m = mesh.rect(12, 6, bottom=mesh.WALL, top=mesh.WALL)
    om = O.OracleMesh(m)
    s = O.OracleSolver(om, O.LAMINAR, mach, 0.0, re)
    y = om.cc[:, 1]
    U = np.column_stack([np.ones(m.nc), np.ones(m.nc), a * y, np.zeros(m.nc)])
    s.set_prim(U)
    s.residual()
    Fv = s.field("Fv_all", 4, "nf")
    mu = np.sqrt(G) * mach / re
    far_cells = set(m.flc[m.frc == mesh.FARFIELD])
    bottom = [f for f in np.nonzero(m.frc == mesh.WALL)[0] if om.fc[f, 1] < 1e-12 and m.flc[f] not in far_cells]
    assert len(bottom) == 10
    for f in bottom:
        assert np.allclose(om.fn[f], [0, -1], atol=1e-15)
        A = om.fa[f]
        assert np.allclose(Fv[f], [0.0, 0.0, -mu * a * A, 0.0], rtol=1e-12, atol=1e-16)


def test_viscous_spectral_radius_and_diagonal_closed_form():
    """Uniform state rho = 2, p = 2 (T = 1, mu = mu_inf) in a box of
    no-slip walls on a uniform dx x dy grid:
      Lambda_v = sum_f max(4/(3 rho), gamma/rho) (mu/Pr) A_f^2 / V
               = (gamma/rho)(mu/Pr) (2 dx^2 + 2 dy^2)/(dx dy)   in every cell,
      Lambda_c = sum_f (|V_f . n| + c) A_f  with V_f = 0 on the walls,
      D = V/dt + 0.75 Lambda_c + Lambda_v with dt = CFL V / Lambda_c
    (ViscousFlow2D.cpp:220-252, TopologyHolderStrategy.cpp:190-215; the
    viscous D adds Lambda_v although dt ignores it, ViscousFlow2D.h:223-242).
    Catches A instead of A^2, a missing /V or /Pr, rho in the numerator, or
    Lambda_v left out of D."""
    mach, re = 0.25, 300.0
    nx, ny, lx, ly = 6, 5, 0.6, 0.4
    m = mesh.rect(nx, ny, lx=lx, ly=ly, bottom=mesh.WALL, top=mesh.WALL, sides=mesh.WALL)
    om = O.OracleMesh(m)
    s = O.OracleSolver(om, O.LAMINAR, mach, 0.0, re)
    u0, v0 = 0.3, 0.1
    U = np.tile([2.0, 2.0, u0, v0], (m.nc, 1))
    s.set_prim(U)
    s.residual()
    dx, dy = lx / nx, ly / ny
    mu = np.sqrt(G) * mach / re
    lamv = s.field("lamv", 1)
    assert np.allclose(lamv, (G / 2.0) * (mu / PR) * (2 * dx * dx + 2 * dy * dy) / (dx * dy), rtol=1e-13, atol=0)
    c = np.sqrt(G * 2.0 / 2.0)
    expect = np.zeros(m.nc)
    for f in range(m.nf):
        vertical = abs(om.fn[f, 1]) < 1e-12
        A = dy if vertical else dx
        vn = 0.0 if m.frc[f] < 0 else (abs(u0) if vertical else abs(v0))
        for cell in (m.flc[f], m.frc[f]):
            if cell >= 0:
                expect[cell] += (vn + c) * A
    lamc = s.field("lamc", 1)
    assert np.allclose(lamc, expect, rtol=1e-13, atol=0)
    s.sweeps()
    D = s.field("D", 4)[:, 0]
    assert np.allclose(D, lamc / 400.0 + 0.75 * lamc + lamv, rtol=1e-13, atol=0)


# --------------------------------------------------------------------------
# SST k-omega (row a14)
# --------------------------------------------------------------------------

def _sst(m, mach=0.15, re=1e7, aoa=0.0):
    om = O.OracleMesh(m)
    return om, O.OracleSolver(om, O.SST, mach, aoa, re)


def _set_sst_state(s, U, k, w, mut=None):
    tU = s.field("tU", 2)
    tU[:, 0] = k
    tU[:, 1] = w
    if mut is not None:
        s.field("mut", 1)[:] = mut
    s.set_prim(U)          # W and (rho k, rho omega) from U and tU


def test_sst_free_stream_decay():
    """A uniform free stream on a far-field box (no walls: F1 = F2 = 0):
    the flow residual vanishes, the turbulence gradients vanish, and the
    turbulence residual is the homogeneous decay, R = (beta* rho k omega,
    beta2 rho omega^2) V.  One LU-SGS iteration then lowers both k and
    omega, as dk/dt = -beta* k omega, domega/dt = -beta2 omega^2 require.
    This pins the source sign of deviation D1: the reference's +Q V
    (SSTKWStrategy.h:993-994) would make R negative and k grow."""
    m = mesh.rect(6, 5, jitter=0.15, seed=2)
    om, s = _sst(m, mach=0.6, aoa=7.0, re=1e5)
    s.residual()
    assert np.abs(s.R).max() < 1e-13
    assert np.abs(s.field("tgrad", 4)).max() < 1e-9 * s.s.fsnd_turb[1]
    assert np.abs(s.field("F1", 1)).max() < 1e-30 and np.abs(s.field("F2", 1)).max() < 1e-15
    k, w = s.s.fsnd_turb[0], s.s.fsnd_turb[1]
    tR = s.field("tR", 2)
    assert np.allclose(tR[:, 0], BSTAR * k * w * om.vol, rtol=1e-10, atol=0)
    assert np.allclose(tR[:, 1], BETA2 * w * w * om.vol, rtol=1e-10, atol=0)
    s.lusgs_step()
    tU = s.field("tU", 2)
    assert np.all(tU[:, 0] < k) and np.all(tU[:, 1] < w)
    assert np.abs(s.U - np.array(list(s.s.fs_primVar)[:4])).max() < 1e-13


def test_sst_blending_functions_closed_form():
    """F1 = tanh(arg1^4), F2 = tanh(arg2^2) with
      arg1 = min(max(sqrt(k)/(beta* omega d), 500 nu/(d^2 omega)), 4 rho sigma_w2 k/(CD d^2)),
      arg2 = max(2 sqrt(k)/(beta* omega d), 500 nu/(d^2 omega)),
      CD = max(2 rho sigma_w2 grad k . grad omega / omega, 1e-20)
    (Menter 1994; SSTKWStrategy.h:695-722), each branch driven in turn by
    hand-set wall distances: the sqrt(k) branch (F1 at arguments 0.2-2),
    the viscous branch, and the cross-diffusion cap (k and omega linear in
    x, so grad k . grad omega is exact), in the cells without a boundary
    face (the far-field faces carry the free stream's k and omega).  Catches a wrong beta*, arg^2 for
    arg^4, nu without 1/rho, or a missing cap."""
    m = mesh.rect(10, 10)
    om, s = _sst(m, mach=0.2, re=2e5)
    nu_scale = np.sqrt(G) * 0.2 / 2e5           # mu at T = 1
    inner, _ = interior_masks(m)
    x = om.cc[:, 0]
    fs = list(s.s.fs_primVar)[:4]
    U = np.tile([1.5, 1.5, fs[2], fs[3]], (m.nc, 1))   # T = 1, rho = 1.5
    rho = 1.5
    nu = nu_scale / rho
    rng = np.random.default_rng(3)

    # (1) sqrt(k) branch: uniform k, omega; d chosen for arguments in [0.2, 2]
    k0, w0 = 0.01, 10.0
    target = rng.uniform(0.2, 2.0, m.nc)
    d = np.sqrt(k0) / (BSTAR * w0 * target)
    om.wdist_wall[:] = d
    _set_sst_state(s, U, k0, w0)
    s.residual()
    assert np.all(500 * nu / (d * d * w0) < 0.5 * target)
    assert np.allclose(s.field("F1", 1)[inner], np.tanh(target[inner] ** 4), rtol=1e-13, atol=1e-16)
    assert np.allclose(s.field("F2", 1)[inner], np.tanh((2 * target[inner]) ** 2), rtol=1e-13, atol=1e-16)

    # (2) viscous branch: k tiny, d chosen for 500 nu/(d^2 omega) in [0.2, 2]
    k0, w0 = 1e-14, 10.0
    d = np.sqrt(500 * nu / (w0 * target))
    om.wdist_wall[:] = d
    _set_sst_state(s, U, k0, w0)
    s.residual()
    assert np.all(np.sqrt(k0) / (BSTAR * w0 * d) < 0.5 * target)
    assert np.allclose(s.field("F1", 1)[inner], np.tanh(target[inner] ** 4), rtol=1e-12, atol=1e-16)
    assert np.allclose(s.field("F2", 1)[inner], np.tanh(target[inner] ** 2), rtol=1e-12, atol=1e-16)

    # (3) the cross-diffusion cap: k = 1 + 50 x, omega = 1 + 2 x
    kx, wx = 50.0, 2.0
    kc, wc = 1.0 + kx * x, 1.0 + wx * x
    cap_target = rng.uniform(0.3, 1.2, m.nc)
    CD = 2 * rho * SIGW2 * kx * wx / wc
    d = np.sqrt(4 * rho * SIGW2 * kc / (CD * cap_target))
    om.wdist_wall[:] = d
    _set_sst_state(s, U, kc, wc)
    s.residual()
    assert np.allclose(s.field("tgrad", 4)[inner], np.tile([kx, 0, wx, 0], (inner.sum(), 1)), rtol=1e-12,
                       atol=1e-10)
    assert np.all((np.sqrt(kc) / (BSTAR * wc * d))[inner] > 5 * cap_target[inner])
    assert np.allclose(s.field("F1", 1)[inner], np.tanh(cap_target[inner] ** 4), rtol=1e-11, atol=1e-16)


def test_sst_eddy_viscosity_closed_form():
    """mu_t = a1 rho k / max(a1 omega, F2 |Omega|) (Menter 1994;
    SSTKWStrategy.h:194-223, 725-780): in uniform flow Omega = 0 and
    mu_t = rho k / omega (the k-omega relation); in a strong shear
    u = a y at wall distances where 0.35 < F2 < 0.9, mu_t = a1 rho k /
    (F2 |a|).  Catches a1
    applied twice, F2 left out, or the vorticity taken as u_y + v_x."""
    m = mesh.rect(8, 8)
    om, s = _sst(m)
    inner, _ = interior_masks(m)
    y = om.cc[:, 1]
    fs = list(s.s.fs_primVar)[:4]
    rho, k0, w0 = 1.2, 0.02, 3.0
    U = np.tile([rho, rho, fs[2], fs[3]], (m.nc, 1))
    _set_sst_state(s, U, k0, w0)
    s.residual()
    s.eddy_viscosity()
    assert np.allclose(s.field("mut", 1), rho * k0 / w0, rtol=1e-13, atol=0)
    a = 10.0                                       # F2 |Omega| >= 3.5 > a1 omega = 0.93
    U[:, 2] = a * y
    U[:, 3] = 0.0
    t = np.random.default_rng(5).uniform(0.6, 1.2, m.nc)
    om.wdist_wall[:] = 2 * np.sqrt(k0) / (BSTAR * w0 * t)   # arg2 = t: F2 = tanh(t^2) in [0.35, 0.89]
    _set_sst_state(s, U, k0, w0)
    s.residual()
    F2 = np.tanh(t * t)
    assert np.allclose(s.field("F2", 1)[inner], F2[inner], rtol=1e-13, atol=0)
    s.eddy_viscosity()
    assert np.allclose(s.field("mut", 1)[inner], A1 * rho * k0 / (F2[inner] * a), rtol=1e-12, atol=0)


def test_sst_production_and_sources_closed_form():
    """Turbulence residual of deep interior cells, R = -Q V + (convection)
    - (diffusion), with Menter's sources
      Q_k     = P - beta* rho k omega,
      Q_omega = C_w rho P / mu_t - beta rho omega^2 + 2 (1 - F1) rho sigma_w2 grad k . grad omega / omega
    (SSTKWStrategy.h:923-1013, source sign of deviation D1):
    * pure shear u = a y: P = mu_t a^2 (Menter; here the reference's form
      agrees with tau_ij du_i/dx_j);
    * pure dilatation u = a x, v = a y: P = -2/3 rho k div u (mu_t small),
      and the convective outflow rho k div u V;
    * pure normal strain u = a x, v = -a y: the reference evaluates the
      strain part as 2/3 (u_x - v_y)^2 mu_t = 8/3 mu_t a^2 (the 2-D
      restriction of the 3-D identity without its (v_y - w_z)^2 and
      (w_z - u_x)^2 terms; tau_ij du_i/dx_j would give 4 mu_t a^2) -- kept
      as the reference writes it;
    * k and omega linear, no flow: destruction and cross diffusion only,
      with F1 = 0 far from walls and F1 = 1 (beta1, no cross diffusion)
      next to them.
    Catches a flipped source sign, a dropped 2/3 rho k div u, (u_x + v_y)
    in the strain term, C_w or beta unblended, or a missing (1 - F1)."""
    m = mesh.rect(10, 10)
    om, s = _sst(m)
    inner, _ = interior_masks(m)
    x, y, V = om.cc[:, 0], om.cc[:, 1], om.vol
    fs = list(s.s.fs_primVar)[:4]
    rho, k0, w0 = 1.0, 0.05, 4.0
    far = 1e11                                     # no wall: F1 = 0, beta2 and C_w2
    om.wdist_wall[:] = far

    def run(u, v, k, w, mut):
        U = np.column_stack([np.full(m.nc, rho), np.full(m.nc, rho), u, v])
        _set_sst_state(s, U, k, w, mut)
        s.residual()
        return s.field("tR", 2).copy(), s.field("F1", 1).copy()

    # pure shear
    a, mut = 2.0, 3e-3
    tR, F1 = run(a * y, np.zeros(m.nc), k0, w0, mut)
    assert np.all(F1 < 1e-30)
    P = mut * a * a
    ek = -(P - BSTAR * rho * k0 * w0) * V
    ew = -(CW2 * rho * P / mut - BETA2 * rho * w0 * w0) * V
    assert np.allclose(tR[inner, 0], ek[inner], rtol=1e-11, atol=0)
    assert np.allclose(tR[inner, 1], ew[inner], rtol=1e-11, atol=0)

    # pure dilatation (convective outflow rho k div u V)
    a, mut = 0.5, 1e-9
    tR, _ = run(a * x, a * y, k0, w0, mut)
    P = -2.0 / 3.0 * rho * k0 * (2 * a)
    ek = -(P - BSTAR * rho * k0 * w0) * V + rho * k0 * 2 * a * V
    ew = -(CW2 * rho * P / mut - BETA2 * rho * w0 * w0) * V + rho * w0 * 2 * a * V
    assert np.allclose(tR[inner, 0], ek[inner], rtol=1e-10, atol=0)
    assert np.allclose(tR[inner, 1], ew[inner], rtol=1e-10, atol=0)

    # pure normal strain (divergence free)
    a, mut = 1.5, 2e-3
    tR, _ = run(a * x, -a * y, k0, w0, mut)
    P = 8.0 / 3.0 * mut * a * a
    ek = -(P - BSTAR * rho * k0 * w0) * V
    assert np.allclose(tR[inner, 0], ek[inner], rtol=1e-11, atol=0)

    # destruction and cross diffusion, no flow; F1 = 0 on the left half, 1 on the right
    kx, wx = 0.04, 3.0
    kc, wc = k0 + kx * x, w0 + wx * x
    near = x > 0.5
    om.wdist_wall[:] = np.where(near, 1e-7, far)
    tR, F1 = run(np.zeros(m.nc), np.zeros(m.nc), kc, wc, 1e-3)
    assert np.all(F1[near] == 1.0) and np.all(F1[~near] < 1e-30)
    # cells whose neighbours share their F1 (the diffusion coefficient is uniform around them)
    same = inner.copy()
    for f in np.nonzero(m.frc >= 0)[0]:
        l, r = m.flc[f], m.frc[f]
        if near[l] != near[r]:
            same[l] = same[r] = False
    assert same[near].sum() > 10 and same[~near].sum() > 10
    ek = BSTAR * rho * kc * wc * V
    cross = 2 * rho * SIGW2 * kx * wx / wc
    ew = np.where(near, BETA1 * rho * wc * wc, BETA2 * rho * wc * wc - cross) * V
    assert np.allclose(tR[same, 0], ek[same], rtol=1e-10, atol=0)
    assert np.allclose(tR[same, 1], ew[same], rtol=1e-10, atol=0)


def test_sst_wall_omega():
    """Wall faces carry k = 0 and Menter's omega_wall = 10 * 6 nu /
    (beta1 d^2) with d the wall cell's wall distance (SSTKWStrategy.h:
    616-659; Menter 1994).  Catches the factor 6 without 10, beta* for
    beta1, or d instead of d^2."""
    mach, re = 0.15, 6e6
    m = mesh.rect(8, 6, ly=0.1, bottom=mesh.WALL, top=mesh.WALL)
    om, s = _sst(m, mach=mach, re=re)
    s.residual()
    Te = s.field("tUedge", 2, "nf")
    mu = np.sqrt(G) * mach / re                  # free-stream T at the no-slip wall
    walls = np.nonzero(m.frc == mesh.WALL)[0]
    d = om.wdist_wall[m.flc[walls]]
    assert np.all(Te[walls, 0] == 0.0)
    assert np.allclose(Te[walls, 1], 60.0 * mu / (1.0 * BETA1 * d * d), rtol=1e-13, atol=0)


# --------------------------------------------------------------------------
# integrated wall force (row f3)
# --------------------------------------------------------------------------

def _wall_cell_faces(m):
    walls = np.nonzero(m.frc == mesh.WALL)[0]
    return walls, m.flc[walls]


def test_force_of_uniform_and_linear_pressure():
    """Euler wall force on a closed body (the cylinder O-grid's polygon):
    a uniform pressure gives zero force; a pressure p = 1 + s.x set at each
    wall face's midpoint gives, by the divergence theorem on the polygon
    (the midpoint rule is exact for a linear p on straight edges),
    F = -A_body s on the body.  At AoA 30 deg, with q = gamma M^2 / 2:
      C_L = (-F_x sin a + F_y cos a) / q, C_D = (F_x cos a + F_y sin a) / q.
    Catches a flipped normal, lift and drag swapped, the rotation by -a, or
    a wrong dynamic pressure."""
    og, _ = mesh.cylinder_ogrid(ni=96, nj=6, dr1=0.02, ratio=1.2, dr_max=0.2)
    om = O.OracleMesh(og)
    mach, aoa = 0.4, 30.0
    s = O.OracleSolver(om, O.EULER, mach, aoa, 1e6)
    walls, cells = _wall_cell_faces(og)
    assert len(set(cells)) == len(cells)
    U = np.tile([1.0, 1.0, 0.0, 0.0], (og.nc, 1))
    s.set_prim(U)
    s.residual()
    cl, cd = s.forces()
    assert abs(cl) < 1e-13 and abs(cd) < 1e-13
    sx, sy = 0.4, 0.7
    U[cells, 1] = 1.0 + sx * om.fc[walls, 0] + sy * om.fc[walls, 1]
    s.set_prim(U)
    s.residual()
    cl, cd = s.forces()
    nodes = og.xy[np.unique(og.fnode[walls].ravel())]
    ang = np.sort(np.arctan2(nodes[:, 1], nodes[:, 0]))
    r = np.hypot(nodes[0, 0], nodes[0, 1])
    area = 0.5 * r * r * np.sum(np.sin(np.diff(np.append(ang, ang[0] + 2 * np.pi))))
    q = 0.5 * G * mach * mach
    a = 3.141593 * aoa / 180
    Fx, Fy = -sx * area, -sy * area
    assert abs(cl - (-Fx * np.sin(a) + Fy * np.cos(a)) / q) < 1e-12
    assert abs(cd - (Fx * np.cos(a) + Fy * np.sin(a)) / q) < 1e-12


def test_viscous_force_subtracts_the_wall_shear():
    """Laminar u = a y over a wall at y = 0 (the only wall): the force is
    the sum over the wall faces of p n A minus the viscous momentum flux,
    whose faces carry exactly -mu a A (pinned above), so the drag of the
    wall is mu a per unit length (the fluid drags the wall along +x).  Catches Fv added instead of
    subtracted, or faces other than walls in the sum."""
    mach, re, a = 0.2, 500.0, 0.8
    m = mesh.rect(40, 4, lx=4.0, ly=0.4, bottom=mesh.WALL, top=mesh.FARFIELD)
    om = O.OracleMesh(m)
    s = O.OracleSolver(om, O.LAMINAR, mach, 0.0, re)
    U = np.column_stack([np.ones(m.nc), np.ones(m.nc), a * om.cc[:, 1], np.zeros(m.nc)])
    s.set_prim(U)
    s.residual()
    cl, cd = s.forces()
    walls, _ = _wall_cell_faces(m)
    Fv = s.field("Fv_all", 4, "nf")
    Ue = s.field("U_edge", 4, "nf")
    fx = np.sum(Ue[walls, 1] * om.fn[walls, 0] * om.fa[walls] - Fv[walls, 2])
    fy = np.sum(Ue[walls, 1] * om.fn[walls, 1] * om.fa[walls] - Fv[walls, 3])
    q = 0.5 * G * mach * mach
    assert abs(cd - fx / q) < 1e-13 and abs(cl - fy / q) < 1e-13
    # every wall face carries mu a A: in the two corner cells the far-field
    # side faces add nothing to du/dy (n_y = 0) and their state has v = 0
    # (the free stream at AoA 0 on inflow, the cell's own on outflow), so
    # dv/dx = 0 there too
    mu = np.sqrt(G) * mach / re
    assert abs(fx - mu * a * 4.0) < 1e-12 * mu * a * 4.0
    assert abs(fy - (-1.0 * 4.0)) < 1e-12      # p n A of p = 1 on the wall, normal (0, -1)


def test_sst_wall_distance_closed_form():
    """The SST model's wall distance (deviation D6) is the Euclidean
    distance to the nearest wall face segment.  On the cylinder O-grid every
    cell centre (the vertex mean) lies on the bisector of one wall edge of
    the inscribed N-gon, at radius (r_j + r_j+1)/2 cos(pi/N), so its distance
    is ((r_j + r_j+1)/2 - R) cos(pi/N); in a flat channel it is the height.
    The reference's projected distance |(x_c - x_f) . n_f| (kept as the
    geometry's wdist) is far smaller than the true distance away from the
    wall, which this pins too."""
    ni, nj, dr1, ratio, dmax = 64, 12, 0.01, 1.3, 0.5
    og, _ = mesh.cylinder_ogrid(ni=ni, nj=nj, dr1=dr1, ratio=ratio, dr_max=dmax)
    om = O.OracleMesh(og)
    r = mesh.ogrid_radii(nj, 0.5, dr1, ratio, dmax)
    rc = np.hypot(om.cc[:, 0], om.cc[:, 1]) / np.cos(np.pi / ni)    # (r_j + r_j+1) / 2
    expect = (rc - 0.5) * np.cos(np.pi / ni)
    assert np.allclose(om.wdist_wall, expect, rtol=1e-12, atol=1e-15)
    assert np.all(np.min(np.abs(rc[:, None] - 0.5 * (r[:-1] + r[1:])[None, :]), axis=1) < 1e-12)
    ch = mesh.channel_bump(ni=30, nj=12, nq=3, seed=5, height=0.0)
    oc = O.OracleMesh(ch)
    y = oc.cc[:, 1]
    assert np.allclose(oc.wdist_wall, np.minimum(y, 1 - y), rtol=0, atol=1e-15)
    far = expect > 0.2
    assert np.median(om.wdist[far] / om.wdist_wall[far]) < 0.2
