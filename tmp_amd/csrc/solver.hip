// gfx950 kernels of the residual-and-update path and the context behind the
// C ABI (include/usg.h).
//
// One thread per cell.  Every face loop is a gather over the cell's face slots
// in ascending face order, which reproduces the reference's scatter order
// (faces visited k = 0..nf-1, each adding to its left and right cell) without
// floating-point atomics.  Kernels per LU-SGS iteration:
//   k_grad    boundary states, Green-Gauss gradients (+grad T, +grad k/omega),
//             Venkatakrishnan limiter, SST F1/F2                   (per cell)
//   k_face    limited reconstruction + HLLC, viscous and SST fluxes and the
//             spectral-radius terms, each face once                 (per face)
//   k_cell    residual and spectral radii = ordered, signed gather of the
//             face terms, SST source, and the LU-SGS diagonal factors (per cell)
//   k_sweep   one launch per wave of the forward and the backward sweep; the
//             backward wave of each cell also writes the updated state into
//             the second buffer (no separate update pass)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/usg.h"
#include "geometry.h"
#include "kernels.h"

namespace usg {

// ---------------------------------------------------------------------------
// SST constants (SSTKWStrategy.h:13-29)
// ---------------------------------------------------------------------------
#define SST_BETA1 0.075
#define SST_BETA2 0.0828
#define SST_SK1 0.85
#define SST_SK2 1.0
#define SST_SW1 0.5
#define SST_SW2 0.856
#define SST_CW1 0.533
#define SST_CW2 0.440
#define SST_A1 0.31
#define SST_BSTAR 0.09

constexpr int kBlock = 256;

template <int FLOW>
struct Flow {
  static constexpr bool visc = FLOW != USG_EULER;
  static constexpr bool sst = FLOW == USG_SST;
};

__device__ __forceinline__ void load4(const double *a, int nc, int c, double *v) {
  v[0] = a[c];
  v[1] = a[nc + c];
  v[2] = a[2 * nc + c];
  v[3] = a[3 * nc + c];
}
__device__ __forceinline__ void store4(double *a, int nc, int c, const double *v) {
  a[c] = v[0];
  a[nc + c] = v[1];
  a[2 * nc + c] = v[2];
  a[3 * nc + c] = v[3];
}

// Face slot decode.
struct Slot {
  int f;      // face id
  bool left;  // the cell is the face's left cell
  int nb;     // neighbour cell or BC code
};
__device__ __forceinline__ bool get_slot(const DevMesh &m, int s, int c, Slot &o) {
  int e = m.cf[s * m.nc + c];
  if (e == kNoSlot) return false;
  o.left = e >= 0;
  o.f = o.left ? e : -e - 1;
  o.nb = m.cnb[s * m.nc + c];
  return true;
}

// ---------------------------------------------------------------------------
// k_bstate: the face state of every boundary face (BoundaryStrategy::
// HandleInvicidBoundary / HandleViscousBoundary, BoundaryStrategy.h:73-335)
// from its cell's current state, for the gradient kernel, which would
// otherwise hold the far-field Riemann state's registers (50 VGPRs, one
// wave per SIMD) for the few cells that have a boundary face.  The slip-wall
// check is counted here.
// ---------------------------------------------------------------------------
template <bool VISC, bool B>
__global__ __launch_bounds__(kBlock) void k_bstate(const DevMesh m, const DevState s, const Consts k_) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= m.nbf) return;
  const int f = m.bface[b], c = m.flc[f], nc = m.nc;
  const Consts k = case_consts<B>(m, k_, c);
  double Ul[4], Ue[4];
  load4p(s.U, nc, c, Ul);
  if (!boundary_state<VISC>(k, m.frc[f], Ul, m.fnx[f], m.fny[f], Ue)) {
    atomicAdd(&s.err[1], 1);
    if constexpr (B) atomicOr(&s.err[2 + m.ccase[c]], 2);
  }
  for (int j = 0; j < 4; j++) s.Ub[(size_t)j * m.nbf + b] = Ue[j];
}

// ---------------------------------------------------------------------------
// k_grad: boundary states + gradients + limiter (+ SST F1/F2)
// Euler2D::updateEdgeValues/solveGradient (Euler2D.cpp:202-240),
// ViscousFlow2D.cpp:174-252, Vankatakrishnan_Limiter.hpp:69-200,
// LaminarFluxStrategy::computeMu (cells), SSTKWStrategy.h:616-722.
// ---------------------------------------------------------------------------
// Mesh arrays by compute type.
template <typename T>
__device__ __forceinline__ const T *fnx_of(const DevMesh &m) { return pickc<T>(m.fnx, m.fnxf); }
template <typename T>
__device__ __forceinline__ const T *fny_of(const DevMesh &m) { return pickc<T>(m.fny, m.fnyf); }
template <typename T>
__device__ __forceinline__ const T *fA_of(const DevMesh &m) { return pickc<T>(m.fA, m.fAf); }

// Offset x_f - x_c of face f from the centre of its left (`left`) or right
// cell, precomputed per face (FP64 differences, rounded once for FP32).
template <typename T>
__device__ __forceinline__ void face_offset(const DevMesh &m, int f, bool left, T &v0, T &v1) {
  const T *d = pickc<T>(m.fdx, m.fdxf);
  const int nf = m.nf;
  v0 = left ? d[f] : d[2 * nf + f];
  v1 = left ? d[nf + f] : d[3 * nf + f];
}

// ---------------------------------------------------------------------------
// k_grad: boundary states + gradients + limiter (+ SST F1/F2)
// Euler2D::updateEdgeValues/solveGradient (Euler2D.cpp:202-240),
// ViscousFlow2D.cpp:174-252, Vankatakrishnan_Limiter.hpp:69-200,
// LaminarFluxStrategy::computeMu (cells), SSTKWStrategy.h:616-722.
// T is the compute type (double; float in the mixed-precision mode, where
// boundary states stay FP64 and the cell viscosity is FP64).
// ---------------------------------------------------------------------------
template <int FLOW, typename T, bool B>
__global__ __launch_bounds__(kBlock) void k_grad(const DevMesh m, const DevState s, const Consts k_,
                                                const int *__restrict__ items, const XSlice sl) {
  using F = Flow<FLOW>;
  int c;
  if (!xcd_item(items, sl, c)) return;
  const Consts k = case_consts<B>(m, k_, c);
  const int nc = m.nc;
  const T *fnx = fnx_of<T>(m), *fny = fny_of<T>(m), *fA = fA_of<T>(m);
  double Ucd[4];
  load4p(s.U, nc, c, Ucd);
  T Uc[4];
  for (int j = 0; j < 4; j++) Uc[j] = T(Ucd[j]);
  const double V = m.vol[c];
  T g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  T gT[2] = {0, 0};
  T tg[4] = {0, 0, 0, 0};
  T umax[4] = {Uc[0], Uc[1], Uc[2], Uc[3]};
  T umin[4] = {Uc[0], Uc[1], Uc[2], Uc[3]};
  T tUc[2] = {0, 0};
  double muc = 0;
  if (F::sst) {
    tUc[0] = T(s.tU[pidx(0, nc, c)]);
    tUc[1] = T(s.tU[pidx(1, nc, c)]);
  }
  // FP32 (mixed) path: face values enter Green-Gauss as differences from the
  // cell value, formed in FP64 -- sum (phi_f - phi_c) n A equals
  // sum phi_f n A on a closed cell (sum n A = 0) without cancelling O(1)
  // terms in single precision.  FP64 path: the reference's sum phi_f n A.
  constexpr bool kDiff = std::is_same<T, float>::value;
  double umaxd[4] = {Ucd[0], Ucd[1], Ucd[2], Ucd[3]};
  double umind[4] = {Ucd[0], Ucd[1], Ucd[2], Ucd[3]};
  const double Tcd = Ucd[1] / Ucd[0];
  for (int sl = 0; sl < 4; sl++) {
    Slot o;
    if (!get_slot(m, sl, c, o)) break;
    // FP64: the slot's own copy of the normal and length (coalesced)
    T n0, n1, A;
    if constexpr (std::is_same<T, double>::value) {
      const size_t sc = (size_t)sl * nc + c;
      n0 = m.snx[sc];
      n1 = m.sny[sc];
      A = m.sA[sc];
    } else {
      n0 = fnx[o.f];
      n1 = fny[o.f];
      A = fA[o.f];
    }
    T Ue[4];
    double Ued[4];
    if (o.nb >= 0) {
      double Unbd[4];
      load4p(s.U, nc, o.nb, Unbd);
      for (int j = 0; j < 4; j++) {
        if constexpr (kDiff) {
          Ued[j] = (Unbd[j] + Ucd[j]) / 2;
          Ue[j] = T((Unbd[j] - Ucd[j]) / 2);
          umaxd[j] = smax(umaxd[j], Unbd[j]);
          umind[j] = smin(umind[j], Unbd[j]);
        } else {
          const T Unb = T(Unbd[j]);
          Ue[j] = o.left ? (Uc[j] + Unb) / 2 : (Unb + Uc[j]) / 2;
          umax[j] = smax(umax[j], Unb);
          umin[j] = smin(umin[j], Unb);
        }
      }
    } else {
      // the face state k_bstate formed for this residual
      const int b = m.bidx[o.f];
      for (int j = 0; j < 4; j++) Ued[j] = s.Ub[(size_t)j * m.nbf + b];
      for (int j = 0; j < 4; j++) Ue[j] = kDiff ? T(Ued[j] - Ucd[j]) : T(Ued[j]);
    }
    // Green-Gauss: + for the left cell, - for the right cell
    if (o.left) {
      for (int j = 0; j < 4; j++) {
        g[2 * j] = g[2 * j] + n0 * (A * Ue[j]);
        g[2 * j + 1] = g[2 * j + 1] + n1 * (A * Ue[j]);
      }
    } else {
      for (int j = 0; j < 4; j++) {
        g[2 * j] = g[2 * j] - n0 * (A * Ue[j]);
        g[2 * j + 1] = g[2 * j + 1] - n1 * (A * Ue[j]);
      }
    }
    if (F::visc) {
      T edgeT = kDiff ? T(Ued[1] / Ued[0] - Tcd) : Ue[1] / Ue[0];
      if (o.left) {
        gT[0] = gT[0] + n0 * (A * edgeT);
        gT[1] = gT[1] + n1 * (A * edgeT);
      } else {
        gT[0] = gT[0] - n0 * (A * edgeT);
        gT[1] = gT[1] - n1 * (A * edgeT);
      }
      if (F::sst) {
        // ComputeFaceTValue (SSTKWStrategy.h:616-659)
        T te[2];
        if constexpr (kDiff) {
          const double tc0 = s.tU[pidx(0, nc, c)], tc1 = s.tU[pidx(1, nc, c)];
          double ted[2];
          if (o.nb >= 0) {
            ted[0] = 0.5 * (s.tU[pidx(0, nc, o.nb)] - tc0);
            ted[1] = 0.5 * (s.tU[pidx(1, nc, o.nb)] - tc1);
          } else if (o.nb == kWallD) {
            const double dis = m.wdist[c];
            ted[0] = 0.0 - tc0;
            ted[1] = 10.0 * 6.0 * sutherland<double>(k, Ued[1] / Ued[0]) / (Ued[0] * SST_BETA1 * dis * dis) - tc1;
          } else if (dot2(Ued[2], Ued[3], m.fnx[o.f], m.fny[o.f]) > 0.0) {
            ted[0] = 0.0;
            ted[1] = 0.0;
          } else {
            ted[0] = k.fsnd_turb[0] - tc0;
            ted[1] = k.fsnd_turb[1] - tc1;
          }
          te[0] = T(ted[0]);
          te[1] = T(ted[1]);
        } else {
          if (o.nb >= 0) {
            T t0 = T(s.tU[pidx(0, nc, o.nb)]), t1 = T(s.tU[pidx(1, nc, o.nb)]);
            te[0] = o.left ? T(0.5) * (tUc[0] + t0) : T(0.5) * (t0 + tUc[0]);
            te[1] = o.left ? T(0.5) * (tUc[1] + t1) : T(0.5) * (t1 + tUc[1]);
          } else if (o.nb == kWallD) {
            T dis = T(m.wdist[c]);
            te[0] = 0.0;
            te[1] = T(10.0 * 6.0) * sutherland<T>(k, edgeT) / (Ue[0] * T(SST_BETA1) * dis * dis);
          } else if (dot2(Ue[2], Ue[3], n0, n1) > T(0.0)) {
            te[0] = tUc[0];
            te[1] = tUc[1];
          } else {
            te[0] = T(k.fsnd_turb[0]);
            te[1] = T(k.fsnd_turb[1]);
          }
        }
        for (int q = 0; q < 2; q++) {
          if (o.left) {
            tg[2 * q] = tg[2 * q] + n0 * (A * te[q]);
            tg[2 * q + 1] = tg[2 * q + 1] + n1 * (A * te[q]);
          } else {
            tg[2 * q] = tg[2 * q] - n0 * (A * te[q]);
            tg[2 * q + 1] = tg[2 * q + 1] - n1 * (A * te[q]);
          }
        }
      }
    }
  }
  const T Vt = T(V);
  T *G = pick<T>(s.grad, s.gradf);
  for (int j = 0; j < 8; j++) g[j] = g[j] / Vt;
  for (int j = 0; j < 4; j++) st2(G, j, nc, c, g[2 * j], g[2 * j + 1]);
  if (F::visc) {
    T *GT = pick<T>(s.gradT, s.gradTf);
    gT[0] = gT[0] / Vt;
    gT[1] = gT[1] / Vt;
    st2(GT, 0, nc, c, gT[0], gT[1]);
    double Tc = Ucd[1] / Ucd[0];
    muc = sutherland<double>(k, Tc);
    s.mu[c] = muc;
  }
  // Venkatakrishnan limiter: eps^2 = 5^3 * V_left (integer 3/ndim == 1)
  T d1max[4], d1min[4], ph[4] = {1.0, 1.0, 1.0, 1.0};
  for (int j = 0; j < 4; j++) {
    if constexpr (kDiff) {
      d1max[j] = T(umaxd[j] - Ucd[j]);
      d1min[j] = T(umind[j] - Ucd[j]);
    } else {
      d1max[j] = umax[j] - Uc[j];
      d1min[j] = umin[j] - Uc[j];
    }
  }
  for (int sl = 0; sl < 4; sl++) {
    Slot o;
    if (!get_slot(m, sl, c, o)) break;
    T v0, v1;
    if constexpr (std::is_same<T, double>::value) {
      v0 = m.sdx[(size_t)sl * nc + c];
      v1 = m.sdy[(size_t)sl * nc + c];
    } else {
      face_offset<T>(m, o.f, o.left, v0, v1);
    }
    T Vl = o.left ? Vt : T(m.vol[o.nb]);
    T eps2 = T(125.0) * Vl;
    for (int j = 0; j < 4; j++) {
      T delta2 = dot2(g[2 * j], g[2 * j + 1], v0, v1);
      ph[j] = smin(ph[j], venkat(d1max[j], d1min[j], delta2, eps2));
    }
  }
  // first order (and overset hole and receptor cells, whose gradients reach
  // into hole cells that hold no solution): the face states are the cell values
  if (k.order == 1 || frozen_cell(m, c))
    for (int j = 0; j < 4; j++) ph[j] = 0.0;
  T *P = pick<T>(s.phi, s.phif);
  st2(P, 0, nc, c, ph[0], ph[1]);
  st2(P, 1, nc, c, ph[2], ph[3]);

  if (F::sst) {
    T *TG = pick<T>(s.tgrad, s.tgradf);
    for (int q = 0; q < 4; q++) tg[q] = tg[q] / Vt;
    st2(TG, 0, nc, c, tg[0], tg[1]);
    st2(TG, 1, nc, c, tg[2], tg[3]);
    // ComputeSSTf1f2 (SSTKWStrategy.h:695-722)
    T gradVal = dot2(tg[0], tg[1], tg[2], tg[3]);
    T rho = Uc[0], dis = T(m.wdist[c]);
    T Kk = tUc[0], w = tUc[1];
    T muL = T(muc);
    T CD_kw = smax(T(2.0) * rho * T(SST_SW2) / w * gradVal, T(1.0e-20));
    T arg1 = smax(sqrt(Kk) / (T(0.09) * w * dis), T(500.0) * muL / (rho * w * dis * dis));
    arg1 = smin(arg1, T(4.0) * rho * T(SST_SW2) * Kk / (CD_kw * dis * dis));
    T arg2 = smax(T(2.0) * sqrt(Kk) / (T(0.09) * w * dis), T(500.0) * muL / (rho * w * dis * dis));
    pick<T>(s.F1, s.F1f)[c] = tanh(arg1 * arg1 * arg1 * arg1);
    pick<T>(s.F2, s.F2f)[c] = tanh(arg2 * arg2);
  }
}

// Limited reconstruction of one side: UL = (grad . (x_f - x_c)) * phi + U
// (Vankatakrishnan_Limiter.hpp:118-125, HLLCFluxStrategy.h:56-65).
template <typename T>
__device__ __forceinline__ void reconstruct(const DevMesh &m, const DevState &s, int cell, int f, bool left,
                                            const T *Ucell, T *out) {
  const int nc = m.nc;
  const T *G = pickc<T>(s.grad, s.gradf);
  const T *P = pickc<T>(s.phi, s.phif);
  T v0, v1;
  face_offset<T>(m, f, left, v0, v1);
  const Pair<T> p01 = ld2(P, 0, nc, cell), p23 = ld2(P, 1, nc, cell);
  const T ph[4] = {p01.x, p01.y, p23.x, p23.y};
  for (int j = 0; j < 4; j++) {
    const Pair<T> g = ld2(G, j, nc, cell);
    T d2 = dot2(g.x, g.y, v0, v1);
    out[j] = d2 * ph[j] + Ucell[j];
  }
}

// ---------------------------------------------------------------------------
// mu/Pr of a cell, laminar plus turbulent for SST (the r_A viscosity of
// SolveViscousFlux, TopologyHolderStrategy.cpp:143-188)
template <bool SST>
__device__ __forceinline__ double cell_mpr(const DevState &s, const Consts &k, int c) {
  double mprc = s.mu[c] / k.pr;
  if (SST) mprc = mprc + s.mut[c] / k.prt;
  return mprc;
}

// max(4/(3 rho), gamma/rho) of the viscous spectral radius
// (ViscousFlow2D.cpp:243-249, TopologyHolderStrategy.cpp:217-323).  For
// rho > 0 the second term is the larger (4/3 against gamma = 1.4, far beyond
// rounding, overflow and underflow included), so only it is divided out;
// any other rho takes the whole expression.  The bits are the same.
template <typename T>
__device__ __forceinline__ T visc_spectral_max(T rho) {
  const T Gamma = T(USG_GAMMA);
  return rho > T(0.0) ? Gamma / rho : smax(T(4) / (T(3.0) * rho), Gamma / rho);
}

// r_A = omega (|V_j . n| + c_j) (+ max(4/(3 rho_j), gamma/rho_j) (mu/Pr)_i /
// |x_j - x_i| if viscous) of cell i receiving from neighbour j with state
// Ua (SolveDF, TopologyHolderStrategy.cpp:217-323)
template <bool VISC>
__device__ __forceinline__ double lusgs_ra(const double *Ua, double n0, double n1, double mprc, double dist) {
  const double Gamma = USG_GAMMA, OMEGAN = 1.5;
  double Vn = dot2(Ua[2], Ua[3], n0, n1);
  double cs = sqrt(Gamma * Ua[1] / Ua[0]);
  double rA = OMEGAN * (fabs(Vn) + cs);
  if (VISC) rA += visc_spectral_max(Ua[0]) * (mprc) / dist;
  return rA;
}

// face_terms: every flux of one face, computed once.  HLLCFluxStrategy::computeFlux
// (HLLCFluxStrategy.cpp:4-171) with the limited reconstruction of
// updateLeftRight (HLLCFluxStrategy.h:35-100); the viscous face flux of
// LaminarFluxStrategy::computeFlux (LaminarFluxStrategy.h:110-216); the SST
// convective minus diffusive face flux (SSTKWStrategy.h:616-659,783-920).
// Outputs per face, through out(q, value): q 0-3 the convective flux, 4 the
// spectral term, 5-8 the viscous flux, 9 the viscous spectral term, 10-11
// the SST flux, all area-weighted and oriented out of the left cell; the
// sweeps' r_A of both sides go to their cell slots.
// No occupancy attribute: held to 96 VGPRs (5 waves per SIMD) the SST face
// kernel spills 5 VGPRs and runs 179 us instead of 173 us per launch on the
// 1 M-cell mesh; at its own 104 VGPRs it runs 4 waves without spilling.
// ---------------------------------------------------------------------------
template <int FLOW, typename T, bool B, typename Out>
__device__ __forceinline__ void face_terms(const DevMesh &m, const DevState &s, const Consts &k_, int f, Out &out) {
  using F = Flow<FLOW>;
  const int nc = m.nc, nf = m.nf;
  const T Gamma = T(USG_GAMMA);
  const int L = m.flc[f], Rr = m.frc[f];
  const Consts k = case_consts<B>(m, k_, L);
  const T n0 = fnx_of<T>(m)[f], n1 = fny_of<T>(m)[f], A = fA_of<T>(m)[f];
  double ULd[4], URd[4] = {0, 0, 0, 0};
  load4p(s.U, nc, L, ULd);
  T UL[4], UR[4], Ue[4], Fx[4];
  for (int j = 0; j < 4; j++) UL[j] = T(ULd[j]);
  const bool inner = Rr >= 0;
  if (inner) {
    load4p(s.U, nc, Rr, URd);
    for (int j = 0; j < 4; j++) UR[j] = T(URd[j]);
    T Ua[4], Ub[4];
    reconstruct<T>(m, s, L, f, true, UL, Ua);
    reconstruct<T>(m, s, Rr, f, false, UR, Ub);
    if (!hllc(Ua, Ub, n0, n1, A, Fx)) {
      atomicAdd(&s.err[0], 1);
      if constexpr (B) atomicOr(&s.err[2 + m.ccase[L]], 1);
    }
  } else {
    // boundary faces in FP64 in every mode
    double Ued[4], Fxd[4];
    boundary_state<F::visc>(k, Rr, ULd, m.fnx[f], m.fny[f], Ued);
    boundary_flux(Rr, Ued, m.fnx[f], m.fny[f], m.fA[f], Fxd);
    for (int j = 0; j < 4; j++) {
      Ue[j] = T(Ued[j]);
      Fx[j] = T(Fxd[j]);
    }
  }
  // face record: flux 0-3, spectral term 4, viscous flux 5-8, viscous
  // spectral term 9, SST 10-11
  for (int j = 0; j < 4; j++) out(j, Fx[j]);
  if (inner) {
    // r_A of the sweeps' off-diagonal terms for each side, from the cell
    // states they would otherwise gather: the left cell receives from the
    // right (U_R, mu/Pr of L) and the right cell from the left
    double mprL = 0.0, mprR = 0.0, dist = 0.0;
    if (F::visc) {
      mprL = cell_mpr<F::sst>(s, k, L);
      mprR = cell_mpr<F::sst>(s, k, Rr);
      dist = m.fdist[f];
    }
    out.ra((size_t)m.fslot[f] * nc + L, lusgs_ra<F::visc>(URd, m.fnx[f], m.fny[f], mprL, dist),
           (size_t)m.fslot[nf + f] * nc + Rr, lusgs_ra<F::visc>(ULd, m.fnx[f], m.fny[f], mprR, dist));
  }
  if (inner)
    for (int j = 0; j < 4; j++) Ue[j] = (UL[j] + UR[j]) / 2;
  {
    // spectral radius term (Euler2D::solveSpectralRadius, Euler2D.cpp:242-266)
    T cs = sqrt(Gamma * Ue[1] / Ue[0]);
    T Vn = dot2(Ue[2], Ue[3], n0, n1);
    out(4, (fabs(Vn) + cs) * A);
  }
  if (!F::visc) return;
  const T *G = pickc<T>(s.grad, s.gradf);
  const T *GT = pickc<T>(s.gradT, s.gradTf);
  T mue = sutherland<T>(k, Ue[1] / Ue[0]);
  T mutf = 0.0;
  if (F::sst) {
    if (inner) {
      mutf = T(0.5) * (T(s.mut[L]) + T(s.mut[Rr]));
    } else if (Rr == kWallD) {
      mutf = 0.0;
    } else {
      T Vn = dot2(Ue[2], Ue[3], n0, n1);
      mutf = Vn > T(0.0) ? T(s.mut[L]) : T(k.fsnd_mut);
    }
  }
  T mu_at_face = F::sst ? mue + mutf : mue;
  T mpr = mue / T(k.pr);
  if (F::sst) mpr = mpr + mutf / T(k.prt);
  T efficient = T(2.0) / T(3.0) * mu_at_face;
  T gE[3][2];
  T l2r0 = 0, l2r1 = 0, l2sq = 0;
  if (inner) {
    l2r0 = pickc<T>(m.fdx, m.fdxf)[4 * nf + f];
    l2r1 = pickc<T>(m.fdx, m.fdxf)[5 * nf + f];
    l2sq = dot2(l2r0, l2r1, l2r0, l2r1);
    for (int j = 0; j < 2; j++) {
      const Pair<T> gl = ld2(G, 2 + j, nc, L), gr = ld2(G, 2 + j, nc, Rr);
      T a0 = (gl.x + gr.x) * T(0.5);
      T a1 = (gl.y + gr.y) * T(0.5);
      // right - left differences are formed in FP64 (exact for T = double)
      T sc = (T(URd[2 + j] - ULd[2 + j]) - dot2(a0, a1, l2r0, l2r1)) / l2sq;
      gE[j][0] = a0 + l2r0 * sc;
      gE[j][1] = a1 + l2r1 * sc;
    }
    const Pair<T> tl2 = ld2(GT, 0, nc, L), tr2 = ld2(GT, 0, nc, Rr);
    T a0 = (tl2.x + tr2.x) * T(0.5);
    T a1 = (tl2.y + tr2.y) * T(0.5);
    double rightT = URd[1] / URd[0];
    double leftT = ULd[1] / ULd[0];
    T sc = (T(rightT - leftT) - dot2(a0, a1, l2r0, l2r1)) / l2sq;
    gE[2][0] = a0 + l2r0 * sc;
    gE[2][1] = a1 + l2r1 * sc;
  } else {
    const Pair<T> gu = ld2(G, 2, nc, L), gv = ld2(G, 3, nc, L);
    gE[0][0] = gu.x;
    gE[0][1] = gu.y;
    gE[1][0] = gv.x;
    gE[1][1] = gv.y;
    if (Rr == kWallD) {
      gE[2][0] = 0;
      gE[2][1] = 0;
    } else {
      const Pair<T> gt = ld2(GT, 0, nc, L);
      gE[2][0] = gt.x;
      gE[2][1] = gt.y;
    }
  }
  T taoxx = efficient * (T(2.0) * gE[0][0] - gE[1][1]);
  T taoyy = efficient * (T(2.0) * gE[1][1] - gE[0][0]);
  T taoxy = mu_at_face * (gE[0][1] + gE[1][0]);
  T qx = -Gamma / (Gamma - 1) * (mpr)*gE[2][0];
  T qy = -Gamma / (Gamma - 1) * (mpr)*gE[2][1];
  T Fv[4];
  Fv[0] = 0.0;
  Fv[1] = ((Ue[2] * taoxx + Ue[3] * taoxy - qx) * n0 + (Ue[2] * taoxy + Ue[3] * taoyy - qy) * n1) * A;
  Fv[2] = (taoxx * n0 + taoxy * n1) * A;
  Fv[3] = (taoxy * n0 + taoyy * n1) * A;
  for (int j = 0; j < 4; j++) out(5 + j, Fv[j]);
  // viscous spectral radius term (ViscousFlow2D.cpp:243-249); each cell
  // divides it by its own volume
  out(9, visc_spectral_max(Ue[0]) * mpr * A * A);
  if (!F::sst) return;
  // face turbulence values (ComputeFaceTValue, SSTKWStrategy.h:616-659)
  const T Vn = dot2(Ue[2], Ue[3], n0, n1);
  T tl[2] = {T(s.tU[pidx(0, nc, L)]), T(s.tU[pidx(1, nc, L)])}, tr[2] = {0, 0}, te[2];
  if (inner) {
    tr[0] = T(s.tU[pidx(0, nc, Rr)]);
    tr[1] = T(s.tU[pidx(1, nc, Rr)]);
    te[0] = T(0.5) * (tl[0] + tr[0]);
    te[1] = T(0.5) * (tl[1] + tr[1]);
  } else if (Rr == kWallD) {
    T dis = T(m.wdist[L]);
    te[0] = 0.0;
    te[1] = T(10.0 * 6.0) * mue / (Ue[0] * T(SST_BETA1) * dis * dis);
  } else if (Vn > T(0.0)) {
    te[0] = tl[0];
    te[1] = tl[1];
  } else {
    te[0] = T(k.fsnd_turb[0]);
    te[1] = T(k.fsnd_turb[1]);
  }
  // convective (SolveConvectiveTFlux :783-831)
  T Fct[2];
  if (inner) {
    T Vp = T(0.5) * (Vn + fabs(Vn));
    T Vs = T(0.5) * (Vn - fabs(Vn));
    for (int q = 0; q < 2; q++) Fct[q] = tl[q] * UL[0] * Vp * A + tr[q] * UR[0] * Vs * A;
  } else {
    for (int q = 0; q < 2; q++) Fct[q] = Vn * Ue[0] * te[q] * A;
  }
  // diffusive (SolveViscousTFlux :834-920)
  const T *F1 = pickc<T>(s.F1, s.F1f);
  const T *TG = pickc<T>(s.tgrad, s.tgradf);
  T face_f1;
  T gq[2][2];
  if (inner) {
    face_f1 = T(0.5) * (F1[L] + F1[Rr]);
    for (int q = 0; q < 2; q++) {
      const Pair<T> ql = ld2(TG, q, nc, L), qr = ld2(TG, q, nc, Rr);
      gq[q][0] = (ql.x + qr.x) * T(0.5);
      gq[q][1] = (ql.y + qr.y) * T(0.5);
    }
    for (int q = 0; q < 2; q++) {
      const T dq = T(s.tU[pidx(q, nc, Rr)] - s.tU[pidx(q, nc, L)]);
      gq[q][0] = gq[q][0] + (dq - (gq[q][0] * l2r0 + gq[q][1] * l2r1)) * l2r0 / l2sq;
      gq[q][1] = gq[q][1] + (dq - (gq[q][0] * l2r0 + gq[q][1] * l2r1)) * l2r1 / l2sq;
    }
  } else {
    face_f1 = F1[L];
    for (int q = 0; q < 2; q++) {
      const Pair<T> ql = ld2(TG, q, nc, L);
      gq[q][0] = ql.x;
      gq[q][1] = ql.y;
    }
  }
  T sigma_k = face_f1 * T(SST_SK1) + (1 - face_f1) * T(SST_SK2);
  T sigma_w = face_f1 * T(SST_SW1) + (1 - face_f1) * T(SST_SW2);
  T txx_k = (mue + sigma_k * mutf) * gq[0][0];
  T tyy_k = (mue + sigma_k * mutf) * gq[0][1];
  T txx_w = (mue + sigma_w * mutf) * gq[1][0];
  T tyy_w = (mue + sigma_w * mutf) * gq[1][1];
  T Fv0 = txx_k * n0 + tyy_k * n1;
  T Fv1 = txx_w * n0 + tyy_w * n1;
  Fv0 *= A;
  Fv1 *= A;
  out(10, Fct[0] - Fv0);
  out(11, Fct[1] - Fv1);
}


// Quantities of the face record by flow: Euler 5, laminar 10, SST 12.
template <int FLOW>
__host__ __device__ constexpr int face_q() { return FLOW == USG_SST ? 12 : (FLOW == USG_LAMINAR ? 10 : 5); }

// The face record keeps its quantities in pairs, [pair][nf][2], so that a
// face's two quantities come in one 16-byte access; the pairs follow the
// order k_cell reads them in: convective flux (0,1) (2,3), viscous flux
// (5,6) (7,8), spectral terms (4,9), SST (10,11).  Euler: (0,1) (2,3) (4,-).
template <int FLOW>
__host__ __device__ constexpr int rec_pairs() { return FLOW == USG_SST ? 6 : (FLOW == USG_LAMINAR ? 5 : 3); }
template <int FLOW>
__host__ __device__ constexpr int rec_pair(int q) {
  return q < 4 ? q >> 1 : (q == 4 || q == 9) ? (FLOW == USG_EULER ? 2 : 4) : q < 9 ? 2 + ((q - 5) >> 1) : 5;
}
__host__ __device__ constexpr int rec_comp(int q) {
  return q < 4 ? (q & 1) : q == 4 ? 0 : q == 9 ? 1 : q < 9 ? ((q - 5) & 1) : q - 10;
}
// The quantity held in component `comp` of pair p (-1: the Euler pad).
template <int FLOW>
__host__ __device__ constexpr int rec_q(int p, int comp) {
  return p < 2 ? 2 * p + comp
         : FLOW == USG_EULER ? (comp == 0 ? 4 : -1)
         : p < 4 ? 5 + 2 * (p - 2) + comp
         : p == 4 ? (comp == 0 ? 4 : 9) : 10 + comp;
}
// The pair of the convective flux, viscous flux, spectral terms and SST flux.
template <int FLOW>
struct RecPairs {
  static constexpr int conv0 = 0, conv1 = 1, visc0 = 2, visc1 = 3;
  static constexpr int lam = FLOW == USG_EULER ? 2 : 4, sst = 5;
};

// The face record held in registers until face_terms has made every load:
// a store issued among the loads would make the compiler reload what it may
// alias and keep later loads behind it.  Then each pair of quantities goes to
// FF[pair][f] as one 16-byte store (contiguous over a wave) and the two r_A
// values to their cell slots.
template <typename T>
struct FaceOutRegs {
  T v[12];
  size_t ia = 0, ib = 0;
  double ra_a = 0.0, ra_b = 0.0;
  bool has_ra = false;
  __device__ __forceinline__ void operator()(int q, T x) { v[q] = x; }
  __device__ __forceinline__ void ra(size_t a, double va, size_t b, double vb) {
    ia = a, ra_a = va, ib = b, ra_b = vb, has_ra = true;
  }
};

template <int FLOW, typename T, bool B>
__global__ __launch_bounds__(kBlock) void k_face(const DevMesh m, const DevState s, const Consts k_,
                                                const int *__restrict__ items, const XSlice sl) {
  int f;
  if (!xcd_item(items, sl, f)) return;
  FaceOutRegs<T> o;
  face_terms<FLOW, T, B>(m, s, k_, f, o);
  using T2 = typename std::conditional<std::is_same<T, double>::value, double2, float2>::type;
  T2 rec[rec_pairs<FLOW>()];
  for (int p = 0; p < rec_pairs<FLOW>(); p++) rec[p].x = rec[p].y = T(0);
  for (int q = 0; q < face_q<FLOW>(); q++) {
    if (rec_comp(q) == 0)
      rec[rec_pair<FLOW>(q)].x = o.v[q];
    else
      rec[rec_pair<FLOW>(q)].y = o.v[q];
  }
  T2 *FF = reinterpret_cast<T2 *>(pick<T>(s.FF, s.FFf));
  for (int p = 0; p < rec_pairs<FLOW>(); p++) FF[(size_t)p * m.nf + f] = rec[p];
  if (o.has_ra) {
    s.rA[o.ia] = o.ra_a;
    s.rA[o.ib] = o.ra_b;
  }
}

// ---------------------------------------------------------------------------
// Per-cell ends of the two sweeps, shared by the sweep waves and by the
// fused first waves (no lower / no upper neighbour: dFi = 0 exactly, so the
// same expressions give the same bits):
//   fwd_store: dW1 = (-R - sum_lower DF) / D        (LUSGS_Integrator.cpp:492-523)
//   bwd_store: dW = dW1 - sum_upper DF / D, W += dW, U = prim(W), SST update
//              (LUSGS_Integrator.cpp:242-255,526-561, SSTKWStrategy.h:194-223)
// Each makes all of its loads before its first store: on gfx950 one counter
// tracks a wave's loads and stores in issue order, so a load issued after a
// store is waited for only together with that store, and the compiler may
// not move a load above a store it cannot prove disjoint.
// ---------------------------------------------------------------------------
// fwd_store from the cell's residual R, tR and SST diagonal tD already in
// registers.
template <int FLOW>
__device__ __forceinline__ void fwd_store_with(const DevMesh &m, const DevState &s, int c, const double *R,
                                               const double *tR, const double *tD, const double *dFi,
                                               const double *tFi, double diag, double *dW1, double *tdW1) {
  using F = Flow<FLOW>;
  const int nc = m.nc;
  // overset hole and receptor cells take no increment (their lower
  // neighbours then see dW1 = 0 across the overset boundary)
  const bool fz = frozen_cell(m, c);
  for (int j = 0; j < 4; j++) dW1[j] = fz ? 0.0 : (-R[j] - dFi[j]) / diag;
  if (F::sst)
    for (int q = 0; q < 2; q++) tdW1[q] = fz ? 0.0 : (-tR[q] - tFi[q]) / tD[q];
  store4p(s.dW1, nc, c, dW1);
  if (F::sst) {
    s.tdW1[pidx(0, nc, c)] = tdW1[0];
    s.tdW1[pidx(1, nc, c)] = tdW1[1];
  }
}

template <int FLOW>
__device__ __forceinline__ void fwd_store(const DevMesh &m, const DevState &s, int c, const double *dFi,
                                          const double *tFi, double diag, double *dW1, double *tdW1) {
  using F = Flow<FLOW>;
  const int nc = m.nc;
  double R[4], tR[2] = {0.0, 0.0}, tD[2] = {1.0, 1.0};
  load4p(s.R, nc, c, R);
  if (F::sst)
    for (int q = 0; q < 2; q++) {
      tR[q] = s.tR[pidx(q, nc, c)];
      tD[q] = s.tD[pidx(q, nc, c)];
    }
  fwd_store_with<FLOW>(m, s, c, R, tR, tD, dFi, tFi, diag, dW1, tdW1);
}

// A cell that is not updated this iteration carries its state into the next
// buffer unchanged (a frozen batch case, an overset hole or receptor cell).
template <int FLOW>
__device__ __forceinline__ void carry_state(const DevMesh &m, const DevState &s, int c) {
  using F = Flow<FLOW>;
  const int nc = m.nc;
  double W[4], U[4], tw[2] = {0.0, 0.0}, tu[2] = {0.0, 0.0}, mt = 0.0;
  load4p(s.W, nc, c, W);
  load4p(s.U, nc, c, U);
  if (F::sst) {
    for (int q = 0; q < 2; q++) {
      tw[q] = s.tW[pidx(q, nc, c)];
      tu[q] = s.tU[pidx(q, nc, c)];
    }
    mt = s.mut[c];
  }
  const double zero[4] = {0.0, 0.0, 0.0, 0.0};
  store4p(s.dW, nc, c, zero);
  store4p(s.Wn, nc, c, W);
  store4p(s.Un, nc, c, U);
  if (F::sst) {
    for (int q = 0; q < 2; q++) {
      s.tdW[pidx(q, nc, c)] = 0.0;
      s.tWn[pidx(q, nc, c)] = tw[q];
      s.tUn[pidx(q, nc, c)] = tu[q];
    }
    s.mutn[c] = mt;
  }
}

template <int FLOW, typename T, bool B>
__device__ __forceinline__ void bwd_store(const DevMesh &m, const DevState &s, const Consts &k, int c,
                                          const double *dW1, const double *tdW1, const double *dFi,
                                          const double *tFi, double diag) {
  using F = Flow<FLOW>;
  const int nc = m.nc;
  if constexpr (B) {
    // a batch case whose residual failed a check in this iteration keeps
    // its state (the other cases are independent of it)
    if (s.err[2 + m.ccase[c]]) {
      carry_state<FLOW>(m, s, c);
      return;
    }
  }
  if (frozen_cell(m, c)) {
    carry_state<FLOW>(m, s, c);
    return;
  }
  double W[4], tD[2] = {1.0, 1.0}, tW[2] = {0.0, 0.0}, g5 = 0.0, g6 = 0.0, f2 = 0.0;
  load4p(s.W, nc, c, W);
  if (F::sst) {
    for (int q = 0; q < 2; q++) {
      tD[q] = s.tD[pidx(q, nc, c)];
      tW[q] = s.tW[pidx(q, nc, c)];
    }
    const T *G = pickc<T>(s.grad, s.gradf);
    g5 = double(G[pidx(5, nc, c)]);
    g6 = double(G[pidx(6, nc, c)]);
    f2 = double(pickc<T>(s.F2, s.F2f)[c]);
  }
  double dW[4], Wn[4], Un[4];
  for (int j = 0; j < 4; j++) {
    dW[j] = dW1[j] - dFi[j] / diag;
    Wn[j] = W[j] + dW[j];
  }
  prim_of(Wn, Un);
  store4p(s.dW, nc, c, dW);
  store4p(s.Wn, nc, c, Wn);
  store4p(s.Un, nc, c, Un);
  if (F::sst) {
    double tdw[2], tw[2];
    for (int q = 0; q < 2; q++) {
      tdw[q] = tdW1[q] - tFi[q] / tD[q];
      s.tdW[pidx(q, nc, c)] = tdw[q];
      tw[q] = tW[q] + tdw[q];
    }
    double rho = Un[0];
    double tu[2] = {tw[0] / rho, tw[1] / rho};
    // floor at 1e-4 of the free stream (the clip kept commented out at
    // SSTKWStrategy.h:206-217); a NaN is floored too
    for (int q = 0; q < 2; q++) {
      double fl = k.fsnd_turb[q] * 1.0e-4;
      if (!(tu[q] >= fl)) {
        tu[q] = fl;
        tw[q] = rho * fl;
      }
    }
    s.tWn[pidx(0, nc, c)] = tw[0];
    s.tWn[pidx(1, nc, c)] = tw[1];
    s.tUn[pidx(0, nc, c)] = tu[0];
    s.tUn[pidx(1, nc, c)] = tu[1];
    double curl = fabs(g6 - g5);
    double h = smax(SST_A1 * tw[1] / rho, f2 * curl);
    s.mutn[c] = SST_A1 * tw[0] / h;
  }
}

// ---------------------------------------------------------------------------
// k_cell: residual of one cell as the reference's scatter order gives it --
// convective face fluxes in ascending face order, then the viscous pass
// (R_L -= Fv, R_R += Fv, LaminarFluxStrategy.h:192-210), SST source then
// faces (SolveTotalTFlux, SSTKWStrategy.h:969-1013, source sign -Q*V), the
// spectral radii (Euler2D.cpp:242-266, ViscousFlow2D.cpp:220-252) -- and
// the per-cell LU-SGS factors: dt (SolveTime, Euler2D.h:67-77 /
// ViscousFlow2D.h:223-242), D (SolveDiag, TopologyHolderStrategy.cpp:190-215),
// the SST diagonal (SSTKWStrategy.h:170-192) and mu/Pr of the cell for r_A.
// ---------------------------------------------------------------------------
template <int FLOW, typename T, bool B, typename In>
__device__ __forceinline__ void cell_terms(const DevMesh &m, const DevState &s, const Consts &k_, int c, In in) {
  using F = Flow<FLOW>;
  const Consts k = case_consts<B>(m, k_, c);
  const int nc = m.nc;
  int fs[4];
  int ns = 0;
  for (int sl = 0; sl < 4; sl++) {
    int e = m.cf[sl * nc + c];
    if (e == kNoSlot) break;
    fs[ns++] = e;
  }
  // faces in ascending order, each added with this cell's sign (left +,
  // right -): the reference's scatter order.  Sums are FP64 in every mode.
  using RP = RecPairs<FLOW>;
  double Rc[4] = {0, 0, 0, 0};
  for (int i = 0; i < ns; i++) {
    const int e = fs[i];
    const double2 a = in.pair(i, e, RP::conv0), b = in.pair(i, e, RP::conv1);
    const double v[4] = {a.x, a.y, b.x, b.y};
    if (e >= 0) {
      for (int j = 0; j < 4; j++) Rc[j] += v[j];
    } else {
      for (int j = 0; j < 4; j++) Rc[j] -= v[j];
    }
  }
  if (F::visc) {
    for (int i = 0; i < ns; i++) {
      const int e = fs[i];
      const double2 a = in.pair(i, e, RP::visc0), b = in.pair(i, e, RP::visc1);
      const double v[4] = {a.x, a.y, b.x, b.y};
      if (e >= 0) {
        for (int j = 0; j < 4; j++) Rc[j] -= v[j];
      } else {
        for (int j = 0; j < 4; j++) Rc[j] += v[j];
      }
    }
  }
  const double V = m.vol[c];
  if (k.dts[0] != 0.0) {
    // dual time stepping: the physical-time derivative of W by backward
    // differences, added after the face fluxes
    for (int j = 0; j < 4; j++)
      Rc[j] += V * (k.dts[0] * s.W[pidx(j, nc, c)] + k.dts[1] * s.Wt1[pidx(j, nc, c)] + k.dts[2] * s.Wt2[pidx(j, nc, c)]);
  }
  double lamc = 0, lamv = 0;
  for (int i = 0; i < ns; i++) {
    const double2 l = in.pair(i, fs[i], RP::lam);
    lamc += l.x;
    if (F::visc) lamv += l.y / V;
  }
  const double OMEGAN = 1.5;
  double dtc;
  if (F::sst)
    dtc = k.cfl * V / (lamc + 4.0 * lamv);
  else
    dtc = k.cfl * V / (lamc);
  double diag = V / dtc + 0.5 * OMEGAN * lamc;
  if (F::visc) diag += lamv;
  if (k.dts[0] != 0.0) diag += k.dts[0] * V;
  // every load of the cell is made before its first store (a store among
  // the loads would hold the later loads behind it)
  double trc[2] = {0.0, 0.0}, tdiag[2] = {0.0, 0.0};
  if (F::sst) {
    double tUc[2] = {s.tU[pidx(0, nc, c)], s.tU[pidx(1, nc, c)]};
    const T *G = pickc<T>(s.grad, s.gradf), *TG = pickc<T>(s.tgrad, s.tgradf);
    double tgc[4];
    for (int q = 0; q < 4; q++) tgc[q] = double(TG[pidx(q, nc, c)]);
    const double F1c = double(pickc<T>(s.F1, s.F1f)[c]), mutc = s.mut[c];
    const double rho = s.U[pidx(0, nc, c)];
    double gux = double(G[pidx(4, nc, c)]), guy = double(G[pidx(5, nc, c)]);
    double gvx = double(G[pidx(6, nc, c)]), gvy = double(G[pidx(7, nc, c)]);
    double ddls0 = 2.0 / 3.0 * ((gux - gvy) * (gux - gvy));
    double ddls1 = (guy + gvx) * (guy + gvx);
    double ddls2 = 2.0 / 3.0 * rho * tUc[0] * (gux + gvy);
    double P = (ddls0 + ddls1) * mutc - ddls2;
    double rdiv = dot2(tgc[0], tgc[1], tgc[2], tgc[3]);
    double Cw = F1c * SST_CW1 + (1.0 - F1c) * SST_CW2;
    double beta = F1c * SST_BETA1 + (1.0 - F1c) * SST_BETA2;
    double kk = tUc[0], w = tUc[1];
    double Q0 = P - SST_BSTAR * rho * kk * w;
    double Q1 = Cw * rho * P / mutc - beta * rho * w * w + 2 * (1 - F1c) * rho * SST_SW2 / w * rdiv;
    trc[0] = -(Q0 * V);
    trc[1] = -(Q1 * V);
    for (int i = 0; i < ns; i++) {
      const int e = fs[i];
      const double2 t = in.pair(i, e, RP::sst);
      if (e >= 0) {
        trc[0] += t.x;
        trc[1] += t.y;
      } else {
        trc[0] -= t.x;
        trc[1] -= t.y;
      }
    }
    // SSTkomegaModel::SolveDiag
    double diagt = V / dtc + 0.5 * OMEGAN * lamc + lamv;
    tdiag[0] = diagt + 2.0 * SST_BSTAR * w * V;
    double gradVal = rdiv;
    tdiag[1] = diagt + V * fabs(2.0 * beta * w + 2.0 * SST_SW2 * (1 - F1c) * gradVal / (w * w));
  }
  const unsigned char wf = m.wflag[c];
  s.lamc[c] = lamc;
  if (F::visc) s.lamv[c] = lamv;
  store4p(s.R, nc, c, Rc);
  s.dt[c] = dtc;
  s.D[c] = diag;
  if (F::sst) {
    s.tR[pidx(0, nc, c)] = trc[0];
    s.tR[pidx(1, nc, c)] = trc[1];
    s.tD[pidx(0, nc, c)] = tdiag[0];
    s.tD[pidx(1, nc, c)] = tdiag[1];
  }
  // cells without lower neighbours (forward wave 0): dW1 = (-R - 0) / D here,
  // and the backward update too for cells with no upper neighbour either
  if (wf & 1) {
    const double zero[4] = {0.0, 0.0, 0.0, 0.0};
    double dW1[4], tdW1[2];
    fwd_store_with<FLOW>(m, s, c, Rc, trc, tdiag, zero, zero, diag, dW1, tdW1);
    if (wf & 2) bwd_store<FLOW, T, B>(m, s, k, c, dW1, tdW1, zero, zero, diag);
  }
}


// The face record in global memory (written by k_face), read by slot, one
// pair of quantities per access.
template <int FLOW, typename T>
struct FaceInGlobal {
  const T *FF;
  int nf;
  __device__ __forceinline__ double2 pair(int, int e, int p) const {
    using T2 = typename std::conditional<std::is_same<T, double>::value, double2, float2>::type;
    const T2 v = reinterpret_cast<const T2 *>(FF)[(size_t)p * nf + (e >= 0 ? e : -e - 1)];
    return make_double2(double(v.x), double(v.y));
  }
};

template <int FLOW, typename T, bool B>
__global__ __launch_bounds__(kBlock) void k_cell(const DevMesh m, const DevState s, const Consts k_,
                                                const int *__restrict__ items, const XSlice sl) {
  int c;
  if (!xcd_item(items, sl, c)) return;
  cell_terms<FLOW, T, B>(m, s, k_, c, FaceInGlobal<FLOW, T>{pickc<T>(s.FF, s.FFf), m.nf});
}

// ---------------------------------------------------------------------------
// k_facecell: face_terms and cell_terms in one launch, the face record held in
// LDS instead of a round trip through HBM (FP64).  Launch block b serves the
// cells k_cell's block b serves (a chunk of the part-grouped Hilbert list:
// spatially compact, of every colour); its tile's faces -- every face of
// those cells, ascending -- are computed first, each into LDS slot
// lds[q*stride + i], then each cell gathers its faces from LDS in ascending
// face order, as k_cell gathers them from the face record.  A face between
// two tiles is computed by both (the same bits), so no tile waits for
// another.  Results equal k_face + k_cell bit for bit.
// ---------------------------------------------------------------------------
struct FaceOutLds {
  double *L;
  int stride, i;
  double *rA;
  __device__ __forceinline__ void operator()(int q, double v) const { L[q * stride + i] = v; }
  __device__ __forceinline__ void ra(size_t a, double va, size_t b, double vb) const {
    rA[a] = va;
    rA[b] = vb;
  }
};
template <int FLOW>
struct FaceInLds {
  const double *L;
  int stride;
  const int *lf;  // local face index of each slot of the cell
  __device__ __forceinline__ double2 pair(int i, int, int p) const {
    const int q0 = rec_q<FLOW>(p, 0), q1 = rec_q<FLOW>(p, 1);
    return make_double2(L[q0 * stride + lf[i]], q1 >= 0 ? L[q1 * stride + lf[i]] : 0.0);
  }
};

template <int FLOW, bool B>
__global__ __launch_bounds__(kBlock) void k_facecell(const DevMesh m, const DevState s, const Consts k_,
                                                    const int *__restrict__ items, const XSlice sl,
                                                    const int *__restrict__ tfoff, const int *__restrict__ tface,
                                                    const unsigned short *__restrict__ lfs, int stride) {
  extern __shared__ double lds[];
  const int t = blockIdx.x, f0 = tfoff[t], f1 = tfoff[t + 1];
  for (int i = f0 + (int)threadIdx.x; i < f1; i += (int)blockDim.x) {
    FaceOutLds o{lds, stride, i - f0, s.rA};
    face_terms<FLOW, double, B>(m, s, k_, tface[i], o);
  }
  __syncthreads();
  int c;
  if (!xcd_item(items, sl, c)) return;
  int lf[4];
  for (int q = 0; q < 4; q++) lf[q] = lfs[(size_t)q * m.nc + c];
  cell_terms<FLOW, double, B>(m, s, k_, c, FaceInLds<FLOW>{lds, stride, lf});
}

// ---------------------------------------------------------------------------
// k_sweep: one wave of the LU-SGS forward (FWD) or backward sweep.
// LUSGSIntegrator::SolveForwardSweep/SolveBackwardSweep (LUSGS_Integrator.cpp:492-561),
// SolveDF/SolveViscousFlux (TopologyHolderStrategy.cpp:143-323); SST
// SolveDF (SSTKWStrategy.h:251-417, deviations D2/D4).  A wave is a slice of
// a part-grouped cell list (a colour block of a colour-blocked numbering, or
// a dependency level).  The backward wave also applies
// UpdatePrimitiveVariable (LUSGS_Integrator.cpp:242-255) into the second
// state buffer, and the SST update (SSTKWStrategy.h:194-223).
// Blocks from nb.nsweep on run residual-norm work instead (NormBlocks).
// ---------------------------------------------------------------------------
__device__ void norm_partial_block(const DevMesh &m, const double *R, const int *b0, const int *b1,
                                   double *partial, int b);
__device__ void norm_ring_block(const double *partial, const int *boff, double *ring, int cap, int *ctr, int cs,
                                int ncase);

// Residual-norm blocks appended to a sweep launch of a single case with
// kBlock-thread blocks: launch blocks [nsweep, nsweep + n) run the norm's
// per-block partial sums (kind 1: n = one per norm block) or its ring sum
// (kind 2: n = 1), the same code on the same cells as k_norm_partial and
// k_norm_ring, which they replace.  R is not written by the sweeps, and the
// ring sum runs in the sweep launch after the partial sums (forward or
// backward).
struct NormBlocks {
  int kind = 0, nsweep = 0, cap = 0;
  const double *R = nullptr;
  const int *b0 = nullptr, *b1 = nullptr, *boff = nullptr;
  double *partial = nullptr, *ring = nullptr;
  int *ctr = nullptr;
};

// DF of one dependency neighbour nb of cell c across slot sl (SolveDF,
// TopologyHolderStrategy.cpp:217-323; SST SolveDF, SSTKWStrategy.h:251-417),
// given nb's increments (dW1 in the forward sweep, dW in the backward one):
// the term t that the cell's sum adds.
template <int FLOW>
__device__ __forceinline__ void df_term(const DevMesh &m, const DevState &s, int c, int sl, bool left, int nb,
                                        const double *dW, const double *tdw, double *t, double *tt) {
  using F = Flow<FLOW>;
  const int nc = m.nc;
  const size_t sc = (size_t)sl * nc + c;
  const double n0 = m.snx[sc], n1 = m.sny[sc], A = m.sA[sc];
  double W[4], Wp[4], Fc[4], Fcp[4];
  for (int j = 0; j < 4; j++) {
    const double W0 = s.W[pidx(j, nc, nb)];
    Wp[j] = W0 + dW[j];
    W[j] = W0;
  }
  cons_flux(W, n0, n1, Fc);
  cons_flux(Wp, n0, n1, Fcp);
  const double rA = s.rA[sc];
  for (int j = 0; j < 4; j++) {
    double DFc = left ? (Fcp[j] - Fc[j]) : (Fc[j] - Fcp[j]);
    double Fv = rA * dW[j];
    t[j] = 0.5 * (DFc - Fv) * A;
  }
  if (F::sst) {
    double u = W[2] / W[0], v = W[3] / W[0];
    double vn = dot2(u, v, n0, n1);
    double up = Wp[2] / Wp[0], vp = Wp[3] / Wp[0];
    double vnp = dot2(up, vp, n0, n1);
    for (int q = 0; q < 2; q++) {
      const double tw = s.tW[pidx(q, nc, nb)];
      double twp = tw + tdw[q];
      double fc = tw * vn, fcp = twp * vnp;
      double DFc = left ? (fcp - fc) : (fc - fcp);
      double Fv = rA * tdw[q];
      tt[q] = 0.5 * (DFc - Fv) * A;
    }
  }
}

// Sum of DF over the dependency neighbours of cell c -- lower ones in the
// forward sweep, upper ones in the backward sweep -- in slot order, from
// their stored increments.
template <int FLOW, bool FWD>
__device__ __forceinline__ void df_sum(const DevMesh &m, const DevState &s, int c, double *dFi, double *tFi) {
  using F = Flow<FLOW>;
  const int nc = m.nc;
  const double *dsrc = FWD ? s.dW1 : s.dW;
  const double *tsrc = FWD ? s.tdW1 : s.tdW;
  for (int sl = 0; sl < 4; sl++) {
    Slot o;
    if (!get_slot(m, sl, c, o)) break;
    if (o.nb < 0) continue;
    if (FWD ? !(o.nb < c) : !(o.nb > c)) continue;
    double dW[4], tdw[2] = {0.0, 0.0}, t[4], tt[2] = {0.0, 0.0};
    for (int j = 0; j < 4; j++) dW[j] = dsrc[pidx(j, nc, o.nb)];
    if (F::sst)
      for (int q = 0; q < 2; q++) tdw[q] = tsrc[pidx(q, nc, o.nb)];
    df_term<FLOW>(m, s, c, sl, o.left, o.nb, dW, tdw, t, tt);
    for (int j = 0; j < 4; j++) dFi[j] += t[j];
    if (F::sst)
      for (int q = 0; q < 2; q++) tFi[q] += tt[q];
  }
}

// The increment a sweep stores for cell u (fwd_store / bwd_store), from the
// stored increments of its dependency neighbours: the same expressions, so
// the same bits.
template <int FLOW, bool FWD, bool B>
__device__ __forceinline__ void increment_of(const DevMesh &m, const DevState &s, int u, double *dW, double *tdw) {
  using F = Flow<FLOW>;
  const int nc = m.nc;
  for (int j = 0; j < 4; j++) dW[j] = 0.0;
  tdw[0] = tdw[1] = 0.0;
  if (!FWD) {
    if constexpr (B) {
      if (s.err[2 + m.ccase[u]]) return;  // a failed batch case carries its state (dW = 0)
    }
  }
  if (frozen_cell(m, u)) return;
  double dFu[4] = {0, 0, 0, 0}, tFu[2] = {0, 0};
  df_sum<FLOW, FWD>(m, s, u, dFu, tFu);
  const double diag = s.D[u];
  if (FWD) {
    for (int j = 0; j < 4; j++) dW[j] = (-s.R[pidx(j, nc, u)] - dFu[j]) / diag;
    if (F::sst)
      for (int q = 0; q < 2; q++) tdw[q] = (-s.tR[pidx(q, nc, u)] - tFu[q]) / s.tD[pidx(q, nc, u)];
  } else {
    for (int j = 0; j < 4; j++) dW[j] = s.dW1[pidx(j, nc, u)] - dFu[j] / diag;
    if (F::sst)
      for (int q = 0; q < 2; q++) tdw[q] = s.tdW1[pidx(q, nc, u)] - tFu[q] / s.tD[pidx(q, nc, u)];
  }
}

// The small last wave of a sweep (the 64 cells of the fourth colour on the
// 1 M-cell airfoil) runs in the launch of the wave before it instead of a
// launch of its own, which would cost a whole launch's latency for a
// handful of cells.  Its blocks come first (a multiple of kXcd, so that the
// wave's own blocks keep their part-to-XCD mapping) and start at once, so
// their longer chains of loads overlap the wave's own work.  A tail cell's
// dependency neighbour in the wave of the same launch (wflag bit 2 forward,
// bit 3 backward) may not have stored its increment yet, so the tail
// recomputes it (increment_of).
struct SweepTail {
  XSlice sl;
  int nblk = 0;
};

// One tail cell with four lanes, lane q on face slot q: each lane finds its
// dependency neighbour's increment (recomputed or loaded) and DF term; the
// terms are then added in slot order, as df_sum adds them, and lane 0
// stores.
template <int FLOW, bool FWD, typename T, bool B>
__device__ __forceinline__ void sweep_tail_cell(const DevMesh &m, const DevState &s, const Consts &k_,
                                                const int *__restrict__ items, const SweepTail &tl) {
  using F = Flow<FLOW>;
  const int b = (int)blockIdx.x, part = b % kXcd;
  const int i = tl.sl.off[part] + (b / kXcd) * ((int)blockDim.x >> 2) + ((int)threadIdx.x >> 2);
  if (i >= tl.sl.off[part + 1]) return;  // the four lanes of a cell leave together
  const int q = (int)threadIdx.x & 3;
  const int c = items ? items[i] : i;
  const int nc = m.nc;
  double t[4] = {0.0, 0.0, 0.0, 0.0}, tt[2] = {0.0, 0.0};
  int has = 0;
  Slot o;
  if (get_slot(m, q, c, o) && o.nb >= 0 && (FWD ? o.nb < c : o.nb > c)) {
    has = 1;
    double dW[4], tdw[2] = {0.0, 0.0};
    if (m.wflag[o.nb] & (FWD ? 4 : 8)) {
      increment_of<FLOW, FWD, B>(m, s, o.nb, dW, tdw);
    } else {
      const double *dsrc = FWD ? s.dW1 : s.dW;
      const double *tsrc = FWD ? s.tdW1 : s.tdW;
      for (int j = 0; j < 4; j++) dW[j] = dsrc[pidx(j, nc, o.nb)];
      if (F::sst)
        for (int r = 0; r < 2; r++) tdw[r] = tsrc[pidx(r, nc, o.nb)];
    }
    df_term<FLOW>(m, s, c, q, o.left, o.nb, dW, tdw, t, tt);
  }
  double dFi[4] = {0, 0, 0, 0}, tFi[2] = {0, 0};
  for (int sl = 0; sl < 4; sl++) {
    const int h = __shfl(has, sl, 4);
    double tk[4], ttk[2];
    for (int j = 0; j < 4; j++) tk[j] = __shfl(t[j], sl, 4);
    for (int r = 0; r < 2; r++) ttk[r] = __shfl(tt[r], sl, 4);
    if (h) {
      for (int j = 0; j < 4; j++) dFi[j] += tk[j];
      if (F::sst)
        for (int r = 0; r < 2; r++) tFi[r] += ttk[r];
    }
  }
  if (q != 0) return;
  const Consts k = case_consts<B>(m, k_, c);
  const double diag = s.D[c];
  const double zero[4] = {0.0, 0.0, 0.0, 0.0};
  if (FWD) {
    const bool last = m.wflag[c] & 2;  // loaded before fwd_store's stores
    double dW1[4], tdW1[2];
    fwd_store<FLOW>(m, s, c, dFi, tFi, diag, dW1, tdW1);
    if (last) bwd_store<FLOW, T, B>(m, s, k, c, dW1, tdW1, zero, zero, diag);
  } else {
    double dW1[4], tdW1[2] = {0.0, 0.0};
    for (int j = 0; j < 4; j++) dW1[j] = s.dW1[pidx(j, nc, c)];
    if (F::sst)
      for (int r = 0; r < 2; r++) tdW1[r] = s.tdW1[pidx(r, nc, c)];
    bwd_store<FLOW, T, B>(m, s, k, c, dW1, tdW1, dFi, tFi, diag);
  }
}

// Held to 96 VGPRs (5 waves per SIMD, as before the tail path existed; the
// inlined recomputation of the tail path would otherwise take 106).
template <int FLOW, bool FWD, typename T, bool B>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_sweep(
    const DevMesh m, const DevState s, const Consts k_, const int *__restrict__ items, const XSlice sl,
    const NormBlocks nb, const SweepTail tl) {
  using F = Flow<FLOW>;
  if (nb.kind && (int)blockIdx.x >= nb.nsweep) {
    if (nb.kind == 1)
      norm_partial_block(m, nb.R, nb.b0, nb.b1, nb.partial, (int)blockIdx.x - nb.nsweep);
    else
      norm_ring_block(nb.partial, nb.boff, nb.ring, nb.cap, nb.ctr, 0, 1);
    return;
  }
  if ((int)blockIdx.x < tl.nblk) {
    sweep_tail_cell<FLOW, FWD, T, B>(m, s, k_, items, tl);
    return;
  }
  int c;
  if (!xcd_item_at(items, sl, (int)blockIdx.x - tl.nblk, c)) return;
  const Consts k = case_consts<B>(m, k_, c);
  const int nc = m.nc;
  const double diag = s.D[c];
  const bool last = FWD && (m.wflag[c] & 2);  // loaded before the first store
  double dFi[4] = {0, 0, 0, 0};
  double tFi[2] = {0, 0};
  df_sum<FLOW, FWD>(m, s, c, dFi, tFi);
  const double zero[4] = {0.0, 0.0, 0.0, 0.0};
  if (FWD) {
    double dW1[4], tdW1[2];
    fwd_store<FLOW>(m, s, c, dFi, tFi, diag, dW1, tdW1);
    // a cell without upper neighbours (backward wave 0) ends here
    if (last) bwd_store<FLOW, T, B>(m, s, k, c, dW1, tdW1, zero, zero, diag);
  } else {
    double dW1[4], tdW1[2] = {0.0, 0.0};
    for (int j = 0; j < 4; j++) dW1[j] = s.dW1[pidx(j, nc, c)];
    if (F::sst)
      for (int q = 0; q < 2; q++) tdW1[q] = s.tdW1[pidx(q, nc, c)];
    bwd_store<FLOW, T, B>(m, s, k, c, dW1, tdW1, dFi, tFi, diag);
  }
}

// RK5 stage (RungeKuttaStrategy.h:39-72): W_now = W0 + alpha*dt*(-R)/V ; U = prim(W_now)
__global__ __launch_bounds__(kBlock) void k_rk_stage(const DevMesh m, const DevState s, double cfl, double alpha,
                                                     int stage, int visc_sst) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = m.nc;
  if (c >= nc) return;
  const double V = m.vol[c];
  if (stage == 0) {
    double dtc = visc_sst ? cfl * V / (s.lamc[c] + 4.0 * s.lamv[c]) : cfl * V / (s.lamc[c]);
    s.dt_rk[c] = dtc;
  }
  const double dtc = s.dt_rk[c];
  double Wn[4], Un[4];
  for (int j = 0; j < 4; j++) Wn[j] = s.W0[pidx(j, nc, c)] + alpha * dtc * (-s.R[pidx(j, nc, c)]) / V;
  prim_of(Wn, Un);
  store4p(s.W, nc, c, Wn);
  store4p(s.U, nc, c, Un);
}

// initializeConsData of the current U into W (and W0)
__global__ __launch_bounds__(kBlock) void k_cons(const DevMesh m, const DevState s, double *Wdst) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = m.nc;
  if (c >= nc) return;
  double U[4], W[4];
  load4p(s.U, nc, c, U);
  cons_of(U, W);
  store4p(Wdst, nc, c, W);
}

// SSTkomegaModel::initializeConsData (SSTKWStrategy.h:97-109)
__global__ __launch_bounds__(kBlock) void k_tcons(const DevMesh m, const DevState s) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = m.nc;
  if (c >= nc) return;
  double rho = s.U[pidx(0, nc, c)];
  s.tW[pidx(0, nc, c)] = s.tU[pidx(0, nc, c)] * rho;
  s.tW[pidx(1, nc, c)] = s.tU[pidx(1, nc, c)] * rho;
}

// Free-stream cold start (Euler2D.cpp:153-171, SSTKWStrategy.h:582-613)
__global__ __launch_bounds__(kBlock) void k_freestream(const DevMesh m, const DevState s, const Consts k_, int sst) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = m.nc;
  if (c >= nc) return;
  const Consts k = m.kc ? m.kc[m.ccase[c]] : k_;
  for (int j = 0; j < 4; j++) s.U[pidx(j, nc, c)] = k.fs[j];
  if (sst) {
    s.tU[pidx(0, nc, c)] = k.fsnd_turb[0];
    s.tU[pidx(1, nc, c)] = k.fsnd_turb[1];
    s.mut[c] = k.fsnd_mut;
  }
}

// ||R_rho / V||^2 partial sums, fixed-order tree per block (deterministic).
// Density-residual norm ||R_rho / V||_2 per case.  Block b sums the squares
// of cells [b0[b], b1[b]) (never two cases in one block); one block per case
// then sums its blocks' partials.
// The block bodies take the block's index explicitly, so that the same code
// also runs as blocks appended to a sweep launch (NormBlocks); they need
// kBlock threads.
__device__ void norm_partial_block(const DevMesh &m, const double *R, const int *b0, const int *b1,
                                   double *partial, int b) {
  __shared__ double sh[kBlock];
  const int c = b0[b] + threadIdx.x;
  double v = 0;
  if (c < b1[b] && !frozen_cell(m, c)) {
    double r = R[pidx(0, m.nc, c)] / m.vol[c];
    v = r * r;
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[b] = sh[0];
}
__global__ __launch_bounds__(kBlock) void k_norm_partial(const DevMesh m, const double *R, const int *b0,
                                                        const int *b1, double *partial) {
  norm_partial_block(m, R, b0, b1, partial, blockIdx.x);
}
__device__ __forceinline__ double case_sum(const double *partial, const int *boff, int cs) {
  __shared__ double sh[kBlock];
  const int i0 = boff[cs], n = boff[cs + 1] - i0;
  double v = 0;
  for (int i = threadIdx.x; i < n; i += kBlock) v = v + partial[i0 + i];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
    __syncthreads();
  }
  return sh[0];
}
// one block per case: out[case] = sqrt(sum)
__global__ __launch_bounds__(kBlock) void k_norm_final(const double *partial, const int *boff, double *out) {
  double v = case_sum(partial, boff, blockIdx.x);
  if (threadIdx.x == 0) out[blockIdx.x] = sqrt(v);
}
// one block per case: ring[(ctr % cap) * ncase + case] = sqrt(sum); the
// counter is advanced by this kernel when there is one case, else by
// k_add_int launched after it (no block may see another's increment)
__device__ void norm_ring_block(const double *partial, const int *boff, double *ring, int cap, int *ctr, int cs,
                                int ncase) {
  double v = case_sum(partial, boff, cs);
  if (threadIdx.x == 0) {
    int i = *ctr;
    ring[(size_t)(i % cap) * ncase + cs] = sqrt(v);
    if (ncase == 1) *ctr = i + 1;
  }
}
__global__ __launch_bounds__(kBlock) void k_norm_ring(const double *partial, const int *boff, double *ring,
                                                      int cap, int *ctr) {
  norm_ring_block(partial, boff, ring, cap, ctr, blockIdx.x, gridDim.x);
}
__global__ void k_add_int(int *p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = *p + v;
}
__global__ void k_set_int(int *p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// Viscous momentum flux tau . n A of wall face f of its cell c, from the
// current state: the cell's Green-Gauss velocity gradient with k_grad's
// arithmetic and the wall face's stress with face_terms' (FP64), so the
// result is what a residual evaluation of this state would put in the face
// record -- without touching the iteration's arrays.
__device__ void wall_shear(const DevMesh &m, const DevState &s, const Consts &k, int f, int c, double &gx,
                           double &gy) {
  const int nc = m.nc;
  double Ucd[4];
  load4p(s.U, nc, c, Ucd);
  double g[4] = {0, 0, 0, 0};  // du/dx, du/dy, dv/dx, dv/dy
  for (int sl = 0; sl < 4; sl++) {
    Slot o;
    if (!get_slot(m, sl, c, o)) break;
    const double n0 = m.fnx[o.f], n1 = m.fny[o.f], A = m.fA[o.f];
    double Ue[4];
    if (o.nb >= 0) {
      double Unbd[4];
      load4p(s.U, nc, o.nb, Unbd);
      for (int j = 2; j < 4; j++) Ue[j] = o.left ? (Ucd[j] + Unbd[j]) / 2 : (Unbd[j] + Ucd[j]) / 2;
    } else {
      boundary_state<true>(k, o.nb, Ucd, m.fnx[o.f], m.fny[o.f], Ue);
    }
    for (int j = 0; j < 2; j++) {
      if (o.left) {
        g[2 * j] = g[2 * j] + n0 * (A * Ue[2 + j]);
        g[2 * j + 1] = g[2 * j + 1] + n1 * (A * Ue[2 + j]);
      } else {
        g[2 * j] = g[2 * j] - n0 * (A * Ue[2 + j]);
        g[2 * j + 1] = g[2 * j + 1] - n1 * (A * Ue[2 + j]);
      }
    }
  }
  const double V = m.vol[c];
  for (int j = 0; j < 4; j++) g[j] = g[j] / V;
  double Ue[4];
  boundary_state<true>(k, kWallD, Ucd, m.fnx[f], m.fny[f], Ue);
  const double mu = sutherland<double>(k, Ue[1] / Ue[0]);
  const double n0 = m.fnx[f], n1 = m.fny[f], A = m.fA[f];
  const double efficient = 2.0 / 3.0 * mu;
  const double taoxx = efficient * (2.0 * g[0] - g[3]);
  const double taoyy = efficient * (2.0 * g[3] - g[0]);
  const double taoxy = mu * (g[1] + g[2]);
  gx = (taoxx * n0 + taoxy * n1) * A;
  gy = (taoxy * n0 + taoyy * n1) * A;
}

// Wall force over each case's wall faces (one block per case: wall faces
// [woff[case], woff[case+1])): p n A with the wall cell's pressure, minus
// the viscous momentum flux for viscous flow, all of the current state.
template <bool VISC>
__global__ __launch_bounds__(kBlock) void k_forces(const DevMesh m, const DevState s, const Consts k_,
                                                   const int *wall_faces, const int *wall_cells, const int *woff,
                                                   double *out2) {
  __shared__ double sx[kBlock], sy[kBlock];
  double fx = 0, fy = 0;
  const int i0 = woff[blockIdx.x], nw = woff[blockIdx.x + 1] - i0;
  for (int i = threadIdx.x; i < nw; i += kBlock) {
    int f = wall_faces[i0 + i], c = wall_cells[i0 + i];
    double p = s.U[pidx(1, m.nc, c)];
    double gx = p * m.fnx[f] * m.fA[f], gy = p * m.fny[f] * m.fA[f];
    if (VISC) {
      const Consts k = m.kc ? m.kc[m.ccase[c]] : k_;
      double vx, vy;
      wall_shear(m, s, k, f, c, vx, vy);
      gx = gx - vx;
      gy = gy - vy;
    }
    fx = fx + gx;
    fy = fy + gy;
  }
  sx[threadIdx.x] = fx;
  sy[threadIdx.x] = fy;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sx[threadIdx.x] = sx[threadIdx.x] + sx[threadIdx.x + w];
      sy[threadIdx.x] = sy[threadIdx.x] + sy[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out2[2 * blockIdx.x] = sx[0];
    out2[2 * blockIdx.x + 1] = sy[0];
  }
}

// Donor-stencil apply (MeshBlock::getInterpolatedSolution + updateSolnData,
// UniTioga/dataUpdate.cpp:6-133): dst[recv[r]][k] = sum_m w[r][m] src[donor[r][m]][k],
// weights in order from 0; ROW (q[i*nvar + k]) or COLUMN (q[k*npts + i],
// the solver's own layout).  Entries with indices outside the arrays are
// skipped, so a bad stencil cannot fault.
__global__ __launch_bounds__(kBlock) void k_stencil(const double *src, double *dst, int nvar, int column, long sp,
                                                    long dp, int nrecv, const int *recv, const int *nw,
                                                    const int *donor, const double *w) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nrecv * nvar) return;
  const int r = (int)(t / nvar), kv = (int)(t - (long)r * nvar);
  const int rc = recv[r];
  if (rc < 0 || rc >= dp) return;
  const int n = min(max(nw[r], 0), 4);
  double acc = 0;
  for (int mm = 0; mm < n; mm++) {
    const int d = donor[4 * r + mm];
    if (d < 0 || d >= sp) continue;
    acc += src[column ? (long)kv * sp + d : (long)d * nvar + kv] * w[4 * r + mm];
  }
  dst[column ? (long)kv * dp + rc : (long)rc * nvar + kv] = acc;
}

// Overset receptor cells from their donors, at the start of an iteration:
// U_r = sum_k w_k U_{d_k} over the receptor's donor cells (ascending), W_r
// = cons(U_r).  The stencil composes the reference's coupling stages --
// field-cell values averaged to the donor mesh's nodes (as Euler2D::outPutCp
// averages, Euler2D.cpp:396-422), the nodal donor stencil
// (MeshBlock::getInterpolatedSolution, UniTioga/dataUpdate.cpp:6-113) and
// the receptor cell as the mean of its vertices (Orchestra.cpp:269-292) --
// into one linear map.  Donors are field cells, which this kernel never
// writes.  Four lanes per receptor, lane j summing variable j over the
// donors in order; the donors' indices, weights and values are loaded 16 at
// a time ahead of the ordered sum, so that a receptor's 9-25 donors cost a
// few memory latencies instead of one chain of dependent loads per donor.
__global__ __launch_bounds__(kBlock) void k_overset(const DevMesh m, const DevState s, int n, const int *recv,
                                                    const int *off, const int *donor, const double *w) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = t >> 2, j = t & 3;
  if (r >= n) return;  // the four lanes of a receptor leave together
  const int nc = m.nc, c = recv[r], i1 = off[r + 1];
  constexpr int kChunk = 16;
  double u = 0.0;
  for (int i0 = off[r]; i0 < i1; i0 += kChunk) {
    int d[kChunk];
    double wk[kChunk], x[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; k++)
      if (i0 + k < i1) {
        d[k] = donor[i0 + k];
        wk[k] = w[i0 + k];
      }
#pragma unroll
    for (int k = 0; k < kChunk; k++)
      if (i0 + k < i1) x[k] = s.U[pidx(j, nc, d[k])];
#pragma unroll
    for (int k = 0; k < kChunk; k++)
      if (i0 + k < i1) u += wk[k] * x[k];
  }
  double U[4], W[4];
  for (int q = 0; q < 4; q++) U[q] = __shfl(u, q, 4);
  cons_of(U, W);
  s.U[pidx(j, nc, c)] = U[j];
  s.W[pidx(j, nc, c)] = W[j];
}

// Face states of the current U (Euler2D::updateEdgeValues through
// BoundaryStrategy::HandleBoundary, BoundaryStrategy.h:73-335): the mean of
// the two cells inside, the wall or far-field state at boundary faces; out
// is [4][nf].  Computed on request for the U_edge getter.
template <bool VISC>
__global__ __launch_bounds__(kBlock) void k_uedge(const DevMesh m, const DevState s, const Consts k_, double *out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = m.nc, nf = m.nf;
  if (f >= nf) return;
  const int L = m.flc[f], R = m.frc[f];
  const Consts k = m.kc ? m.kc[m.ccase[L]] : k_;
  double UL[4], Ue[4];
  load4p(s.U, nc, L, UL);
  if (R >= 0) {
    double UR[4];
    load4p(s.U, nc, R, UR);
    for (int j = 0; j < 4; j++) Ue[j] = (UL[j] + UR[j]) / 2;
  } else {
    boundary_state<VISC>(k, R, UL, m.fnx[f], m.fny[f], Ue);
  }
  for (int j = 0; j < 4; j++) out[(size_t)j * nf + f] = Ue[j];
}

// The SST model's wall distance of every cell: the Euclidean distance from
// the cell centre to the nearest wall face segment of its case, seg[4*i ..]
// = (x0, y0, x1, y1) of wall face i, faces [woff[case], woff[case+1]).  The
// reference's projected distance (UnstructBlock.h:265-298) is small near
// the tangent line of any wall face, far from the wall too; its SST, never
// executed there, would blend most of the far field to the near-wall
// branch.
__global__ __launch_bounds__(kBlock) void k_wall_distance(int nc, const double *ccx, const double *ccy,
                                                          const int *ccase, const int *woff, const double *seg,
                                                          double *out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int cs = ccase ? ccase[c] : 0;
  const double cx = ccx[c], cy = ccy[c];
  double d = 100000000000.0;
  for (int i = woff[cs]; i < woff[cs + 1]; i++) {
    const double x0 = seg[4 * i], y0 = seg[4 * i + 1], x1 = seg[4 * i + 2], y1 = seg[4 * i + 3];
    double ax = x1 - x0, ay = y1 - y0;
    double px = cx - x0, py = cy - y0;
    double t = (px * ax + py * ay) / (ax * ax + ay * ay);
    if (!(t > 0.0)) t = 0.0;
    if (t > 1.0) t = 1.0;
    double qx = px - t * ax, qy = py - t * ay;
    d = smin(d, sqrt(qx * qx + qy * qy));
  }
  out[c] = d;
}

// Plain copy of n doubles: the reference kernel for calibrating traffic counters.
__global__ __launch_bounds__(kBlock) void k_copy(double *dst, const double *src, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace usg

// ===========================================================================
// Host context
// ===========================================================================
using namespace usg;
static_assert(kXcd == kParts, "one spatial part per XCD");

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

namespace usg {
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace usg

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return fail(USG_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
};

struct usg_ctx {
  HostMesh hm;
  usg_params prm{};
  Consts k{};
  DevMesh dm{};
  DevState ds{};
  hipStream_t own = nullptr, stream = nullptr;
  std::vector<void *> allocs;
  int *d_fwd = nullptr, *d_bwd = nullptr;  // part-grouped wave lists
  int *d_cell_list = nullptr, *d_face_list = nullptr;
  int *d_wall_faces = nullptr, *d_wall_cells = nullptr;
  double *d_partial = nullptr, *d_norms = nullptr, *d_force = nullptr;
  int nblk = 0;  // blocks of a one-thread-per-cell launch
  int nnb = 0;   // norm blocks
  int norms_cap = 0;
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  size_t ev_used = 0;
  double fs_aoa_rad = 0;
  bool have_residual = false;
  // captured two-iteration LU-SGS graphs, one per starting state buffer
  bool use_graph = true;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  double *gstart[2] = {nullptr, nullptr};
  hipStream_t gstream[2] = {nullptr, nullptr};
  double *d_ring = nullptr;
  int *d_ctr = nullptr;
  // batch: cases, their parameters and constants; norm blocks (cells
  // [nb0[b], nb1[b]) of one case each, per-case block offsets nboff)
  int ncase = 1;
  std::vector<usg_params> cprm;
  std::vector<Consts> ck;
  std::vector<double> aoa_rad;
  int *d_nb0 = nullptr, *d_nb1 = nullptr, *d_nboff = nullptr, *d_woff = nullptr;
  // threads per block of the kernel groups (USG_K_GRAD .. USG_K_SWEEP)
  int blk[4] = {kBlock, kBlock, kBlock, kBlock};
  // overset coupling: receptor cells and their donor stencil (CSR), applied
  // at the start of every LU-SGS iteration; the blanking is dm.blank
  int ov_n = 0;
  int *d_ov_recv = nullptr, *d_ov_off = nullptr, *d_ov_donor = nullptr;
  double *d_ov_w = nullptr;
  signed char *d_blank = nullptr;
  // dual time stepping: physical step, order, physical steps taken
  double dt_phys = 0.0;
  int dt_order = 0, dt_steps = 0;
  // fused face-cell launch (k_facecell, FP64): tiles of tile_bs cells, their
  // faces (offsets, ids), the local face index of every cell slot, and the
  // LDS row length (the most faces of a tile)
  double *d_uedge = nullptr;  // face states for the U_edge getter, [4][nf]
  bool fused = false;
  // the small last wave of the forward / backward sweep runs as a tail of
  // the wave before it (SweepTail)
  bool fwd_tail = false, bwd_tail = false;
  // the face launch and the sweep waves may take their items as indices (a
  // null list): every face once in any order (single case), and sweep waves
  // that are contiguous cell ranges (the colour-blocked numbering), whose
  // parts are then contiguous sub-ranges
  bool face_ident = false, sweep_ident = false;
  int tile_bs = 0, tile_stride = 0;
  int *d_tfoff = nullptr, *d_tface = nullptr;
  unsigned short *d_lfs = nullptr;
};

constexpr int kRing = 4096;

template <typename T>
static int dalloc(usg_ctx *c, T **p, size_t n) {
  void *q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
  HIPCHK(hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(T), c->stream));
  c->allocs.push_back(q);
  *p = (T *)q;
  return USG_OK;
}

template <typename T>
static int upload(usg_ctx *c, T **p, const T *src, size_t n) {
  int rc = dalloc(c, p, n);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(*p, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return USG_OK;
}

// set_fs_variable2 + update_nondim_variable (TopologyHolderStrategy.cpp:325-396)
// with WrapSolver's density 1.225 and pressure 101325 (SolverWrapper.h:55).
static void freestream(const usg_params &prm, Consts &kk, double &aoa_rad) {
  const double R = 287.0, Gamma = 1.4;
  double fs_mach = prm.mach;
  double fs_density = 1.225, fs_pressure = 101325, fs_Re = prm.reynolds;
  double fs_soundSpeed = std::sqrt(Gamma * fs_pressure / fs_density);
  double fs_Temperature = fs_pressure / (R * fs_density);
  double fs_mu = 1.458e-6 * std::pow(fs_Temperature, 1.5) / (fs_Temperature + 110.5);
  double fs_AOA;
  freestream_state(prm.mach, prm.aoa_deg, kk.fs, &fs_AOA);
  kk.fs_c = std::sqrt(Gamma);
  kk.sqrt_gamma = std::sqrt(Gamma);
  kk.mach = fs_mach;
  kk.re = fs_Re;
  kk.cc = 110.5 / fs_Temperature;
  kk.pr = 0.72;
  kk.prt = 0.9;
  kk.cfl = prm.cfl_lusgs;
  kk.order = 2;
  aoa_rad = fs_AOA;
  // SSTkomegaModel::InitializeturbulenceVariable (SSTKWStrategy.h:582-613)
  const double L = 1.0;
  double fs_mut = fs_mu * 0.035;
  kk.fsnd_mut = fs_mut / fs_mu * std::sqrt(Gamma) * fs_mach / fs_Re;
  double t0 = 1.5 * fs_mach * fs_soundSpeed * 0.01 * fs_mach * fs_soundSpeed * 0.01;
  double t1 = fs_density * t0 / (0.035 * fs_mu);
  kk.fsnd_turb[0] = Gamma * t0 / (fs_soundSpeed * fs_soundSpeed);
  kk.fsnd_turb[1] = std::sqrt(Gamma) * L * t1 / fs_soundSpeed;
}

static int grid(int n) { return (n + kBlock - 1) / kBlock; }

// quantities per face: flux 4 + spectral 1 (+ viscous flux 4 + viscous
// spectral 1) (+ SST 2)
static int rec_pairs_of(int flow) {
  return flow == USG_SST ? rec_pairs<USG_SST>() : flow == USG_LAMINAR ? rec_pairs<USG_LAMINAR>() : rec_pairs<USG_EULER>();
}

static XSlice slice(const int *off) {
  XSlice sl;
  for (int p = 0; p <= kXcd; p++) sl.off[p] = off[p];
  return sl;
}
// grid of a part-split launch of bs-thread blocks: kXcd x (blocks of the largest part)
static int xgrid(const XSlice &sl, int bs) {
  int mx = 0;
  for (int p = 0; p < kXcd; p++) mx = std::max(mx, sl.off[p + 1] - sl.off[p]);
  return kXcd * std::max(1, (mx + bs - 1) / bs);
}


constexpr size_t kMaxTileLds = 64 * 1024;

// The fused launch's tiles for blocks of bs cells: tile t is the cells k_cell's
// block t serves; its faces are every face of those cells, ascending.  False
// when a tile's face record would not fit the LDS a block may take.
static bool build_tiles(usg_ctx *c, int bs) {
  if (c->tile_bs == bs && c->d_tfoff) return true;
  const HostMesh &h = c->hm;
  const int nc = h.nc;
  const XSlice cs = slice(h.cell_off.data());
  const int ntile = xgrid(cs, bs);
  std::vector<int> tfoff(ntile + 1, 0), tface, faces;
  std::vector<unsigned short> lfs(4 * (size_t)nc, 0);
  int maxf = 0;
  for (int t = 0; t < ntile; t++) {
    const int p = t % kXcd, ch = t / kXcd;
    const int a = cs.off[p] + ch * bs, b = std::min(a + bs, cs.off[p + 1]);
    faces.clear();
    for (int i = a; i < b; i++) {
      const int cell = h.cell_list[i];
      for (int q = 0; q < h.cnf[cell]; q++) faces.push_back(h.cface[4 * cell + q]);
    }
    std::sort(faces.begin(), faces.end());
    faces.erase(std::unique(faces.begin(), faces.end()), faces.end());
    if (faces.size() > 65535) return false;
    for (int i = a; i < b; i++) {
      const int cell = h.cell_list[i];
      for (int q = 0; q < h.cnf[cell]; q++)
        lfs[(size_t)q * nc + cell] =
            (unsigned short)(std::lower_bound(faces.begin(), faces.end(), h.cface[4 * cell + q]) - faces.begin());
    }
    tface.insert(tface.end(), faces.begin(), faces.end());
    tfoff[t + 1] = (int)tface.size();
    maxf = std::max(maxf, (int)faces.size());
  }
  const int nq = c->prm.flow == USG_SST ? 12 : (c->prm.flow == USG_LAMINAR ? 10 : 5);
  if ((size_t)nq * std::max(maxf, 1) * sizeof(double) > kMaxTileLds) return false;
  for (void *q : {(void *)c->d_tfoff, (void *)c->d_tface, (void *)c->d_lfs}) {
    if (!q) continue;
    auto it = std::find(c->allocs.begin(), c->allocs.end(), q);
    if (it != c->allocs.end()) c->allocs.erase(it);
    (void)hipFree(q);
  }
  c->d_tfoff = c->d_tface = nullptr;
  c->d_lfs = nullptr;
  if (upload(c, &c->d_tfoff, tfoff.data(), tfoff.size()) ||
      upload(c, &c->d_tface, tface.data(), std::max<size_t>(tface.size(), 1)) ||
      upload(c, &c->d_lfs, lfs.data(), lfs.size()))
    return false;
  c->tile_bs = bs;
  c->tile_stride = std::max(maxf, 1);
  return hipStreamSynchronize(c->stream) == hipSuccess;
}

// The fused face-cell launch is taken for FP64 when its tiles fit.  Tiles
// are (re)built eagerly, never while the stream is being captured.
static bool fused_ready(usg_ctx *c) {
  if (!c->fused || c->prm.precision != USG_FP64) return false;
  if (c->tile_bs == c->blk[USG_K_CELL] && c->d_tfoff) return true;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(c->stream, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return false;
  if (build_tiles(c, c->blk[USG_K_CELL])) return true;
  c->fused = false;
  return false;
}

template <int FLOW, typename T, bool B>
static void launch_residual(usg_ctx *c, hipEvent_t ea, hipEvent_t eb) {
  const XSlice cs = slice(c->hm.cell_off.data()), fs = slice(c->hm.face_off.data());
  const int bg = c->blk[USG_K_GRAD], bf = c->blk[USG_K_FACE], bc = c->blk[USG_K_CELL];
  if (c->dm.nbf > 0)
    hipLaunchKernelGGL((k_bstate<Flow<FLOW>::visc, B>), dim3(grid(c->dm.nbf)), dim3(kBlock), 0, c->stream, c->dm,
                       c->ds, c->k);
  hipLaunchKernelGGL((k_grad<FLOW, T, B>), dim3(xgrid(cs, bg)), dim3(bg), 0, c->stream, c->dm, c->ds, c->k,
                     c->d_cell_list, cs);
  if constexpr (std::is_same<T, double>::value) {
    if (fused_ready(c)) {
      if (ea) hipEventRecord(ea, c->stream);
      const size_t lds = (size_t)face_q<FLOW>() * c->tile_stride * sizeof(double);
      hipLaunchKernelGGL((k_facecell<FLOW, B>), dim3(xgrid(cs, bc)), dim3(bc), lds, c->stream, c->dm, c->ds, c->k,
                         c->d_cell_list, cs, c->d_tfoff, c->d_tface, c->d_lfs, c->tile_stride);
      if (eb) hipEventRecord(eb, c->stream);
      return;
    }
  }
  if (ea) hipEventRecord(ea, c->stream);
  hipLaunchKernelGGL((k_face<FLOW, T, B>), dim3(xgrid(fs, bf)), dim3(bf), 0, c->stream, c->dm, c->ds, c->k,
                     c->face_ident ? nullptr : c->d_face_list, fs);
  if (eb) hipEventRecord(eb, c->stream);
  hipLaunchKernelGGL((k_cell<FLOW, T, B>), dim3(xgrid(cs, bc)), dim3(bc), 0, c->stream, c->dm, c->ds, c->k,
                     c->d_cell_list, cs);
}

template <typename T, bool B>
static void launch_residual_t(usg_ctx *c, hipEvent_t ea, hipEvent_t eb) {
  switch (c->prm.flow) {
    case USG_EULER: launch_residual<USG_EULER, T, B>(c, ea, eb); break;
    case USG_LAMINAR: launch_residual<USG_LAMINAR, T, B>(c, ea, eb); break;
    default: launch_residual<USG_SST, T, B>(c, ea, eb); break;
  }
}

static void launch_residual_events(usg_ctx *c, hipEvent_t ea, hipEvent_t eb) {
  const bool mixed = c->prm.precision == USG_MIXED, batch = c->ncase > 1;
  if (mixed)
    batch ? launch_residual_t<float, true>(c, ea, eb) : launch_residual_t<float, false>(c, ea, eb);
  else
    batch ? launch_residual_t<double, true>(c, ea, eb) : launch_residual_t<double, false>(c, ea, eb);
}

// A pooled event pair for one timed launch (nullptr pair when the pool cannot grow).
static std::pair<hipEvent_t, hipEvent_t> next_events(usg_ctx *c) {
  if (c->ev_used == c->ev.size()) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      if (a) hipEventDestroy(a);
      return {nullptr, nullptr};
    }
    c->ev.push_back({a, b});
  }
  return c->ev[c->ev_used++];
}

// Eager residual; in timing mode the flux kernel is bracketed by a pooled event pair.
static void launch_residual_any(usg_ctx *c) {
  if (!c->timing) {
    launch_residual_events(c, nullptr, nullptr);
    return;
  }
  auto e = next_events(c);
  launch_residual_events(c, e.first, e.second);
}

// A sweep's last wave runs as a tail of the wave before it when it holds
// at most kTailMax cells and at least min_waves waves exist (3: wave 0 runs
// inside k_cell or the forward waves, so the tail needs a launch of its own
// wave to ride on).
constexpr int kTailMax = 8192;
static bool sweep_tail(const std::vector<int> &poff, int min_waves) {
  const int nw = (int)poff.size() / (kXcd + 1);
  if (nw < min_waves) return false;
  const int *p = &poff[(size_t)(nw - 1) * (kXcd + 1)];
  return p[kXcd] - p[0] <= kTailMax;
}

// The last wave's items as a tail of the wave before it (SweepTail).
static SweepTail tail_of(const std::vector<int> &poff, int bs) {
  const int nw = (int)poff.size() / (kXcd + 1);
  SweepTail tl;
  tl.sl = slice(&poff[(size_t)(nw - 1) * (kXcd + 1)]);
  tl.nblk = xgrid(tl.sl, bs / 4);  // four lanes per cell
  return tl;
}

// Sweep launches per iteration: each sweep's waves after wave 0, less the
// small last wave when it runs as a tail.
static int sweep_launches(const usg_ctx *c) {
  const int nfw = (int)c->hm.fwd_poff.size() / (kXcd + 1), nbw = (int)c->hm.bwd_poff.size() / (kXcd + 1);
  return std::max(0, nfw - 1 - (c->fwd_tail ? 1 : 0)) + std::max(0, nbw - 1 - (c->bwd_tail ? 1 : 0));
}

// With fuse_norm, the first two sweep launches of the iteration also carry
// the residual norm into the ring (NormBlocks): the partial sums on the
// first, the ring sum on the second (forward or backward: neither writes R).
template <int FLOW, typename T, bool B>
static void launch_sweeps(usg_ctx *c, bool fuse_norm) {
  const HostMesh &h = c->hm;
  const int bs = c->blk[USG_K_SWEEP];
  int launch = 0;
  auto norm_blocks = [&](int nsweep, int &extra) {
    NormBlocks nb;
    extra = 0;
    if (fuse_norm && launch < 2) {
      nb.kind = launch + 1;
      nb.nsweep = nsweep;
      nb.cap = kRing;
      nb.R = c->ds.R;
      nb.b0 = c->d_nb0;
      nb.b1 = c->d_nb1;
      nb.boff = c->d_nboff;
      nb.partial = c->d_partial;
      nb.ring = c->d_ring;
      nb.ctr = c->d_ctr;
      extra = nb.kind == 1 ? c->nnb : 1;
    }
    launch++;
    return nb;
  };
  // wave 0 of each sweep is fused into k_cell / the forward waves (wflag);
  // a small last wave runs as a tail of the wave before it
  const int nw = (int)h.fwd_poff.size() / (kXcd + 1);
  for (int w = 1; w < nw - (c->fwd_tail ? 1 : 0); w++) {
    XSlice sl = slice(&h.fwd_poff[(size_t)w * (kXcd + 1)]);
    const int nmain = xgrid(sl, bs);
    const SweepTail tl = c->fwd_tail && w == nw - 2 ? tail_of(h.fwd_poff, bs) : SweepTail{};
    const int ntail = tl.nblk;
    int extra = 0;
    const NormBlocks nb = norm_blocks(nmain + ntail, extra);
    hipLaunchKernelGGL((k_sweep<FLOW, true, T, B>), dim3(nmain + ntail + extra), dim3(bs), 0, c->stream, c->dm,
                       c->ds, c->k, c->sweep_ident ? nullptr : c->d_fwd, sl, nb, tl);
  }
  const int nbw = (int)h.bwd_poff.size() / (kXcd + 1);
  for (int w = 1; w < nbw - (c->bwd_tail ? 1 : 0); w++) {
    XSlice sl = slice(&h.bwd_poff[(size_t)w * (kXcd + 1)]);
    const int nmain = xgrid(sl, bs);
    const SweepTail tl = c->bwd_tail && w == nbw - 2 ? tail_of(h.bwd_poff, bs) : SweepTail{};
    const int ntail = tl.nblk;
    int extra = 0;
    const NormBlocks nb = norm_blocks(nmain + ntail, extra);
    hipLaunchKernelGGL((k_sweep<FLOW, false, T, B>), dim3(nmain + ntail + extra), dim3(bs), 0, c->stream, c->dm,
                       c->ds, c->k, c->sweep_ident ? nullptr : c->d_bwd, sl, nb, tl);
  }
}

template <typename T, bool B>
static void launch_sweeps_t(usg_ctx *c, bool fuse_norm) {
  switch (c->prm.flow) {
    case USG_EULER: launch_sweeps<USG_EULER, T, B>(c, fuse_norm); break;
    case USG_LAMINAR: launch_sweeps<USG_LAMINAR, T, B>(c, fuse_norm); break;
    default: launch_sweeps<USG_SST, T, B>(c, fuse_norm); break;
  }
}

// per-case norms into dst[0 .. ncase)
static void launch_norm(usg_ctx *c, double *dst) {
  hipLaunchKernelGGL(k_norm_partial, dim3(c->nnb), dim3(kBlock), 0, c->stream, c->dm, c->ds.R, c->d_nb0, c->d_nb1,
                     c->d_partial);
  hipLaunchKernelGGL(k_norm_final, dim3(c->ncase), dim3(kBlock), 0, c->stream, c->d_partial, c->d_nboff, dst);
}

static double *stbuf(usg_ctx *c) { return c->ds.U; }

static void swap_state(usg_ctx *c) {
  std::swap(c->ds.U, c->ds.Un);
  std::swap(c->ds.W, c->ds.Wn);
  if (c->prm.flow == USG_SST) {
    std::swap(c->ds.tU, c->ds.tUn);
    std::swap(c->ds.tW, c->ds.tWn);
    std::swap(c->ds.mut, c->ds.mutn);
  }
}

static int ensure_norms(usg_ctx *c, int n) {
  if (n <= c->norms_cap) return USG_OK;
  if (c->d_norms) {
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(c->d_norms));
    c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void *)c->d_norms));
  }
  int rc = dalloc(c, &c->d_norms, n);
  if (rc) return rc;
  c->norms_cap = n;
  return USG_OK;
}

// Invariant checks of the last launches.  A single case fails the call with
// the reference's message.  In a batch, failed cases are frozen by the
// sweep kernels and reported through usg_case_status; the call succeeds.
static int check_err(usg_ctx *c) {
  std::vector<int> h(2 + (c->ncase > 1 ? c->ncase : 0), 0);
  HIPCHK(hipMemcpyAsync(h.data(), c->ds.err, h.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (h[0] || h[1]) {
    HIPCHK(hipMemsetAsync(c->ds.err, 0, 2 * sizeof(int), c->stream));
    if (c->ncase > 1) {
      HIPCHK(hipStreamSynchronize(c->stream));
      return USG_OK;
    }
    return fail(USG_ERR_NUMERIC, h[0] ? "SOmething wrong at p star (HLLC |p*_L - p*_R| >= 1e-10)"
                                      : "Something wrong with Vector computation (slip wall |t.n| > 1e-7)");
  }
  return USG_OK;
}

static int clear_case_status(usg_ctx *c) {
  if (c->ncase > 1) HIPCHK(hipMemsetAsync(c->ds.err + 2, 0, c->ncase * sizeof(int), c->stream));
  return USG_OK;
}

// A paired device array ([k/2][nc][2], pidx) as [nc][nv] (FP64 or FP32).
template <typename T>
static int pairs_to_aos(usg_ctx *c, const T *dev, int nv, double *out, long count, bool has) {
  const int nc = c->hm.nc;
  if (count != (long)nv * nc) return fail(USG_ERR_ARG, "count mismatch");
  if (!has || !dev) return fail(USG_ERR_ARG, "field not available for this flow type");
  std::vector<T> tmp((size_t)nv * nc);
  HIPCHK(hipMemcpyAsync(tmp.data(), dev, tmp.size() * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < nc; i++)
    for (int k = 0; k < nv; k++) out[(size_t)i * nv + k] = (double)tmp[pidx(k, nc, i)];
  return USG_OK;
}

static void drop_graphs(usg_ctx *c);

extern "C" {

const char *usg_last_error(void) { return g_err.c_str(); }

int usg_ctx_destroy(usg_ctx *c) {
  if (!c) return USG_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  drop_graphs(c);
  for (auto &e : c->ev) {
    hipEventDestroy(e.first);
    hipEventDestroy(e.second);
  }
  for (void *p : c->allocs) hipFree(p);
  if (c->own) hipStreamDestroy(c->own);
  delete c;
  return USG_OK;
}

static int check_params(const usg_params *params) {
  if (params->flow < USG_EULER || params->flow > USG_SST) return fail(USG_ERR_ARG, "Undefined Flow Type");
  if (params->precision != USG_FP64 && params->precision != USG_MIXED) return fail(USG_ERR_ARG, "bad precision");
  return USG_OK;
}

static int ctx_setup(usg_ctx *c, usg_ctx **out);

// The kernels form structure-of-arrays offsets var*nc + cell (var < 8) and
// q*nf + face (q < 6, the FP32 face offsets) in 32-bit arithmetic.
static int check_size(long long nc, long long nf) {
  if (nc * 8 > 2147483647LL || nf * 6 > 2147483647LL)
    return fail(USG_ERR_ARG, "mesh exceeds 32-bit array offsets (at most 268435455 cells and 357913941 faces)");
  return USG_OK;
}

int usg_ctx_create(const usg_mesh_desc *mesh, const usg_params *params, usg_ctx **out) {
  if (!mesh || !params || !out) return fail(USG_ERR_ARG, "null argument");
  *out = nullptr;
  int rc = check_params(params);
  if (rc) return rc;
  if ((rc = check_size(mesh->nc, mesh->nf))) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(USG_ERR_NODEVICE, "no GPU visible");
  usg_ctx *c = new usg_ctx();
  c->prm = *params;
  c->cprm = {*params};
  std::string e = build_mesh(c->hm, mesh->nn, mesh->xy, mesh->nf, mesh->fnode, mesh->flc, mesh->frc, mesh->nc);
  if (!e.empty()) {
    delete c;
    return fail(USG_ERR_MESH, e);
  }
  return ctx_setup(c, out);
}

int usg_batch_create(const usg_mesh_desc *meshes, const usg_params *params, int ncases, usg_ctx **out) {
  if (!meshes || !params || !out || ncases <= 0) return fail(USG_ERR_ARG, "null argument or no cases");
  *out = nullptr;
  for (int i = 0; i < ncases; i++) {
    int rc = check_params(&params[i]);
    if (rc) return rc;
    if (params[i].flow != params[0].flow || params[i].precision != params[0].precision ||
        params[i].cfl_lusgs != params[0].cfl_lusgs || params[i].cfl_rk != params[0].cfl_rk)
      return fail(USG_ERR_ARG, "cases of a batch share flow type, precision and CFL numbers");
  }
  long long total_nc = 0, total_nf = 0;
  for (int i = 0; i < ncases; i++) {
    total_nc += meshes[i].nc;
    total_nf += meshes[i].nf;
  }
  if (int rc = check_size(total_nc, total_nf)) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(USG_ERR_NODEVICE, "no GPU visible");
  std::vector<HostMesh> cases(ncases);
  for (int i = 0; i < ncases; i++) {
    const usg_mesh_desc &d = meshes[i];
    std::string e = build_mesh(cases[i], d.nn, d.xy, d.nf, d.fnode, d.flc, d.frc, d.nc);
    if (!e.empty()) return fail(USG_ERR_MESH, "case " + std::to_string(i) + ": " + e);
  }
  usg_ctx *c = new usg_ctx();
  c->prm = params[0];
  c->cprm.assign(params, params + ncases);
  merge_meshes(cases, c->hm);
  return ctx_setup(c, out);
}

// Device buffers and constants of a context whose host mesh and parameters are set.
static int ctx_setup(usg_ctx *c, usg_ctx **out) {
  const usg_params *params = &c->prm;
  c->ncase = (int)c->cprm.size();
  c->ck.resize(c->ncase);
  c->aoa_rad.resize(c->ncase);
  for (int i = 0; i < c->ncase; i++) freestream(c->cprm[i], c->ck[i], c->aoa_rad[i]);
  c->k = c->ck[0];
  c->fs_aoa_rad = c->aoa_rad[0];
  int rc = USG_OK;
#define CK(x)                  \
  do {                         \
    rc = (x);                  \
    if (rc) {                  \
      usg_ctx_destroy(c);      \
      return rc;               \
    }                          \
  } while (0)
  if (hipStreamCreate(&c->own) != hipSuccess) {
    delete c;
    return fail(USG_ERR_HIP, "hipStreamCreate");
  }
  c->stream = c->own;
  const HostMesh &h = c->hm;
  const int nc = h.nc, nf = h.nf;
  c->nblk = grid(nc);
  std::vector<double> ccx(nc), ccy(nc), fnx(nf), fny(nf), fcx(nf), fcy(nf);
  for (int i = 0; i < nc; i++) {
    ccx[i] = h.cc[2 * i];
    ccy[i] = h.cc[2 * i + 1];
  }
  for (int f = 0; f < nf; f++) {
    fnx[f] = h.fn[2 * f];
    fny[f] = h.fn[2 * f + 1];
    fcx[f] = h.fc[2 * f];
    fcy[f] = h.fc[2 * f + 1];
  }
  std::vector<int> cf(4 * (size_t)nc, kNoSlot), cnb(4 * (size_t)nc, kNoSlot), fslot(2 * (size_t)nf, 0);
  std::vector<double> snx(4 * (size_t)nc, 0.0), sny(4 * (size_t)nc, 0.0), sA(4 * (size_t)nc, 0.0);
  std::vector<double> sdx(4 * (size_t)nc, 0.0), sdy(4 * (size_t)nc, 0.0);
  for (int i = 0; i < nc; i++)
    for (int s = 0; s < h.cnf[i]; s++) {
      int f = h.cface[4 * i + s];
      bool left = h.flc[f] == i;
      cf[(size_t)s * nc + i] = left ? f : -(f + 1);
      cnb[(size_t)s * nc + i] = left ? h.frc[f] : h.flc[f];
      fslot[(left ? (size_t)0 : (size_t)nf) + f] = s;
      snx[(size_t)s * nc + i] = h.fn[2 * f];
      sny[(size_t)s * nc + i] = h.fn[2 * f + 1];
      sA[(size_t)s * nc + i] = h.fa[f];
      sdx[(size_t)s * nc + i] = fcx[f] - ccx[i];
      sdy[(size_t)s * nc + i] = fcy[f] - ccy[i];
    }
  std::vector<double> fd(6 * (size_t)nf, 0.0);
  for (int f = 0; f < nf; f++) {
    const int L = h.flc[f], R = h.frc[f];
    fd[f] = fcx[f] - ccx[L];
    fd[nf + f] = fcy[f] - ccy[L];
    if (R >= 0) {
      fd[2 * (size_t)nf + f] = fcx[f] - ccx[R];
      fd[3 * (size_t)nf + f] = fcy[f] - ccy[R];
      fd[4 * (size_t)nf + f] = ccx[R] - ccx[L];
      fd[5 * (size_t)nf + f] = ccy[R] - ccy[L];
    }
  }
  DevMesh &dm = c->dm;
  dm.nc = nc;
  dm.nf = nf;
  CK(upload(c, (double **)&dm.vol, h.vol.data(), nc));
  CK(upload(c, (double **)&dm.ccx, ccx.data(), nc));
  CK(upload(c, (double **)&dm.ccy, ccy.data(), nc));
  if (params->flow != USG_SST) CK(upload(c, (double **)&dm.wdist, h.wdist.data(), nc));
  CK(upload(c, (int **)&dm.cf, cf.data(), cf.size()));
  CK(upload(c, (int **)&dm.cnb, cnb.data(), cnb.size()));
  CK(upload(c, (int **)&dm.fslot, fslot.data(), fslot.size()));
  CK(upload(c, (double **)&dm.snx, snx.data(), snx.size()));
  CK(upload(c, (double **)&dm.sny, sny.data(), sny.size()));
  CK(upload(c, (double **)&dm.sA, sA.data(), sA.size()));
  CK(upload(c, (double **)&dm.sdx, sdx.data(), sdx.size()));
  CK(upload(c, (double **)&dm.sdy, sdy.data(), sdy.size()));
  CK(upload(c, (int **)&dm.flc, h.flc.data(), nf));
  CK(upload(c, (int **)&dm.frc, h.frc.data(), nf));
  CK(upload(c, (double **)&dm.fnx, fnx.data(), nf));
  CK(upload(c, (double **)&dm.fny, fny.data(), nf));
  CK(upload(c, (double **)&dm.fA, h.fa.data(), nf));
  CK(upload(c, (double **)&dm.fdx, fd.data(), fd.size()));
  {
    // bit 0: forward wave 0 (no lower neighbour), bit 1: backward wave 0 (no
    // upper neighbour); bit 2 (3): the cell lies in the last-but-one forward
    // (backward) wave and the small last wave runs in that wave's launch
    std::vector<unsigned char> wflag(nc, 0);
    if (!h.fwd_poff.empty())
      for (int i = h.fwd_poff[0]; i < h.fwd_poff[kXcd]; i++) wflag[h.fwd_list[i]] |= 1;
    if (!h.bwd_poff.empty())
      for (int i = h.bwd_poff[0]; i < h.bwd_poff[kXcd]; i++) wflag[h.bwd_list[i]] |= 2;
    c->fwd_tail = sweep_tail(h.fwd_poff, 3);
    c->bwd_tail = sweep_tail(h.bwd_poff, 3);
    c->face_ident = c->ncase == 1;
    c->sweep_ident = c->ncase == 1 && h.use_ranges;
    for (int d = 0; d < 2; d++) {
      const std::vector<int> &poff = d ? h.bwd_poff : h.fwd_poff;
      const std::vector<int> &list = d ? h.bwd_list : h.fwd_list;
      if (!(d ? c->bwd_tail : c->fwd_tail)) continue;
      const int w = (int)poff.size() / (kXcd + 1) - 2;
      for (int i = poff[(size_t)w * (kXcd + 1)]; i < poff[(size_t)w * (kXcd + 1) + kXcd]; i++)
        wflag[list[i]] |= d ? 8 : 4;
    }
    CK(upload(c, (unsigned char **)&dm.wflag, wflag.data(), wflag.size()));
  }
  CK(upload(c, &c->d_fwd, h.fwd_list.data(), nc));
  CK(upload(c, &c->d_bwd, h.bwd_list.data(), nc));
  CK(upload(c, &c->d_cell_list, h.cell_list.data(), nc));
  CK(upload(c, &c->d_face_list, h.face_list.data(), nf));
  {
    // boundary faces for k_bstate, and each face's index among them
    std::vector<int> bface, bidx(nf, -1);
    for (int f = 0; f < nf; f++)
      if (h.frc[f] < 0) {
        bidx[f] = (int)bface.size();
        bface.push_back(f);
      }
    dm.nbf = (int)bface.size();
    CK(upload(c, (int **)&dm.bface, bface.data(), std::max<size_t>(bface.size(), 1)));
    CK(upload(c, (int **)&dm.bidx, bidx.data(), nf));
    CK(dalloc(c, &c->ds.Ub, 4 * std::max<size_t>(bface.size(), 1)));
  }
  std::vector<int> wc(h.wall_faces.size());
  for (size_t i = 0; i < wc.size(); i++) wc[i] = h.flc[h.wall_faces[i]];
  CK(upload(c, &c->d_wall_faces, h.wall_faces.data(), h.wall_faces.size()));
  CK(upload(c, &c->d_wall_cells, wc.data(), wc.size()));
  DevState &s = c->ds;
  const bool visc = params->flow != USG_EULER, sst = params->flow == USG_SST;
  const bool mixed = params->precision == USG_MIXED;
  if (mixed) {
    // FP32 face geometry; the offsets are the FP64 differences, rounded once
    std::vector<float> nxf(nf), nyf(nf), af(nf), fdf(fd.size());
    for (int f = 0; f < nf; f++) {
      nxf[f] = (float)fnx[f];
      nyf[f] = (float)fny[f];
      af[f] = (float)h.fa[f];
    }
    for (size_t i = 0; i < fd.size(); i++) fdf[i] = (float)fd[i];
    CK(upload(c, (float **)&dm.fnxf, nxf.data(), nf));
    CK(upload(c, (float **)&dm.fnyf, nyf.data(), nf));
    CK(upload(c, (float **)&dm.fAf, af.data(), nf));
    CK(upload(c, (float **)&dm.fdxf, fdf.data(), fdf.size()));
  }
  CK(dalloc(c, &s.U, 4 * (size_t)nc));
  CK(dalloc(c, &s.W, 4 * (size_t)nc));
  CK(dalloc(c, &s.Un, 4 * (size_t)nc));
  CK(dalloc(c, &s.Wn, 4 * (size_t)nc));
  if (mixed) {
    CK(dalloc(c, &s.gradf, 8 * (size_t)nc));
    CK(dalloc(c, &s.phif, 4 * (size_t)nc));
    CK(dalloc(c, &s.FFf, (size_t)rec_pairs_of(params->flow) * 2 * nf));
  } else {
    CK(dalloc(c, &s.grad, 8 * (size_t)nc));
    CK(dalloc(c, &s.phi, 4 * (size_t)nc));
    CK(dalloc(c, &s.FF, (size_t)rec_pairs_of(params->flow) * 2 * nf));
  }
  CK(dalloc(c, &s.lamc, nc));
  CK(dalloc(c, &s.R, 4 * (size_t)nc));
  CK(dalloc(c, &s.dW1, 4 * (size_t)nc));
  CK(dalloc(c, &s.dW, 4 * (size_t)nc));
  CK(dalloc(c, &s.dt, nc));
  CK(dalloc(c, &s.dt_rk, nc));
  CK(dalloc(c, &s.W0, 4 * (size_t)nc));
  CK(dalloc(c, &s.D, nc));
  CK(dalloc(c, &s.err, 2 + (c->ncase > 1 ? (size_t)c->ncase : 0)));
  if (visc) {
    if (mixed)
      CK(dalloc(c, &s.gradTf, 2 * (size_t)nc));
    else
      CK(dalloc(c, &s.gradT, 2 * (size_t)nc));
    CK(dalloc(c, &s.lamv, nc));
    CK(dalloc(c, &s.mu, nc));
    std::vector<double> dist(nf, 0.0);
    for (int f = 0; f < nf; f++) {
      const int L = h.flc[f], R = h.frc[f];
      if (R < 0) continue;
      double l0 = ccx[R] - ccx[L], l1 = ccy[R] - ccy[L];
      double d2 = 0;
      d2 += l0 * l0;
      d2 += l1 * l1;
      dist[f] = std::sqrt(d2);
    }
    CK(upload(c, (double **)&dm.fdist, dist.data(), nf));
  }
  CK(dalloc(c, &s.rA, 4 * (size_t)nc));
  if (sst) {
    CK(dalloc(c, &s.tU, 2 * (size_t)nc));
    CK(dalloc(c, &s.tW, 2 * (size_t)nc));
    CK(dalloc(c, &s.tUn, 2 * (size_t)nc));
    CK(dalloc(c, &s.tWn, 2 * (size_t)nc));
    if (mixed) {
      CK(dalloc(c, &s.tgradf, 4 * (size_t)nc));
      CK(dalloc(c, &s.F1f, nc));
      CK(dalloc(c, &s.F2f, nc));
    } else {
      CK(dalloc(c, &s.tgrad, 4 * (size_t)nc));
      CK(dalloc(c, &s.F1, nc));
      CK(dalloc(c, &s.F2, nc));
    }
    CK(dalloc(c, &s.mut, nc));
    CK(dalloc(c, &s.mutn, nc));
    CK(dalloc(c, &s.tR, 2 * (size_t)nc));
    CK(dalloc(c, &s.tdW1, 2 * (size_t)nc));
    CK(dalloc(c, &s.tdW, 2 * (size_t)nc));
    CK(dalloc(c, &s.tD, 2 * (size_t)nc));
  }
  {
    // norm blocks: kBlock cells of one case each; per-case block offsets
    std::vector<int> nb0, nb1, nboff(1, 0);
    for (int i = 0; i < c->ncase; i++) {
      for (int a = h.case_cell[i]; a < h.case_cell[i + 1]; a += kBlock) {
        nb0.push_back(a);
        nb1.push_back(std::min(a + kBlock, h.case_cell[i + 1]));
      }
      nboff.push_back((int)nb0.size());
    }
    c->nnb = (int)nb0.size();
    CK(upload(c, &c->d_nb0, nb0.data(), nb0.size()));
    CK(upload(c, &c->d_nb1, nb1.data(), nb1.size()));
    CK(upload(c, &c->d_nboff, nboff.data(), nboff.size()));
    CK(upload(c, &c->d_woff, h.case_wall.data(), h.case_wall.size()));
    if (c->ncase > 1) {
      std::vector<int> ccase(nc);
      for (int i = 0; i < c->ncase; i++)
        for (int a = h.case_cell[i]; a < h.case_cell[i + 1]; a++) ccase[a] = i;
      CK(upload(c, (int **)&dm.ccase, ccase.data(), ccase.size()));
      CK(upload(c, (Consts **)&dm.kc, c->ck.data(), c->ck.size()));
    }
  }
  if (params->flow == USG_SST) {
    // the SST model's wall distance, computed here from the wall segments
    std::vector<double> seg(4 * std::max<size_t>(h.wall_faces.size(), 1), 0.0);
    for (size_t i = 0; i < h.wall_faces.size(); i++) {
      const int f = h.wall_faces[i];
      for (int q = 0; q < 2; q++) {
        seg[4 * i + 2 * q] = h.xy[2 * (size_t)h.fnode[2 * f + q]];
        seg[4 * i + 2 * q + 1] = h.xy[2 * (size_t)h.fnode[2 * f + q] + 1];
      }
    }
    double *dseg = nullptr, *dwd = nullptr;
    CK(upload(c, &dseg, seg.data(), seg.size()));
    CK(dalloc(c, &dwd, nc));
    hipLaunchKernelGGL(k_wall_distance, dim3(grid(nc)), dim3(kBlock), 0, c->stream, nc, dm.ccx, dm.ccy, dm.ccase,
                       c->d_woff, dseg, dwd);
    if (hipGetLastError() != hipSuccess) CK(fail(USG_ERR_HIP, "wall distance launch"));
    dm.wdist = dwd;
  }
  CK(dalloc(c, &c->d_partial, c->nnb));
  CK(dalloc(c, &c->d_force, 2 * (size_t)c->ncase));
  CK(ensure_norms(c, 16 * c->ncase));
  CK(dalloc(c, &c->d_ring, (size_t)kRing * c->ncase));
  CK(dalloc(c, &c->d_ctr, 1));
#undef CK
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    usg_ctx_destroy(c);
    return fail(USG_ERR_HIP, "stream sync after upload");
  }
  *out = c;
  return usg_init_freestream(c);
}

int usg_set_stream(usg_ctx *c, void *st) {
  if (!c) return fail(USG_ERR_ARG, "null ctx");
  c->stream = st ? (hipStream_t)st : c->own;
  return USG_OK;
}

int usg_sync(usg_ctx *c) {
  if (!c) return fail(USG_ERR_ARG, "null ctx");
  HIPCHK(hipStreamSynchronize(c->stream));
  return USG_OK;
}

int usg_init_freestream(usg_ctx *c) {
  if (!c) return fail(USG_ERR_ARG, "null ctx");
  int rc = clear_case_status(c);
  if (rc) return rc;
  const bool sst = c->prm.flow == USG_SST;
  hipLaunchKernelGGL(k_freestream, dim3(c->nblk), dim3(kBlock), 0, c->stream, c->dm, c->ds, c->k, (int)sst);
  hipLaunchKernelGGL(k_cons, dim3(c->nblk), dim3(kBlock), 0, c->stream, c->dm, c->ds, c->ds.W);
  if (sst) hipLaunchKernelGGL(k_tcons, dim3(c->nblk), dim3(kBlock), 0, c->stream, c->dm, c->ds);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->have_residual = false;
  return USG_OK;
}

int usg_set_prim(usg_ctx *c, const double *U, const double *turb) {
  if (!c || !U) return fail(USG_ERR_ARG, "null argument");
  int rc = clear_case_status(c);
  if (rc) return rc;
  const int nc = c->hm.nc;
  std::vector<double> soa(4 * (size_t)nc);
  for (int i = 0; i < nc; i++)
    for (int j = 0; j < 4; j++) soa[pidx(j, nc, i)] = U[4 * (size_t)i + j];
  HIPCHK(hipMemcpyAsync(c->ds.U, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_cons, dim3(c->nblk), dim3(kBlock), 0, c->stream, c->dm, c->ds, c->ds.W);
  if (c->prm.flow == USG_SST) {
    // (k, omega) as given, or as they are; (rho k, rho omega) from them and
    // the new density either way (SSTkomegaModel::initializeConsData)
    std::vector<double> t(turb ? 2 * (size_t)nc : 0);
    if (turb) {
      for (int i = 0; i < nc; i++)
        for (int j = 0; j < 2; j++) t[pidx(j, nc, i)] = turb[2 * (size_t)i + j];
      HIPCHK(hipMemcpyAsync(c->ds.tU, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    hipLaunchKernelGGL(k_tcons, dim3(c->nblk), dim3(kBlock), 0, c->stream, c->dm, c->ds);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->have_residual = false;
  return USG_OK;
}

// [nc][width] host values into a device array, paired ([width/2][nc][2],
// pidx) or one array per component.
static int upload_aos(usg_ctx *c, const double *src, int width, double *dst, bool paired) {
  const int nc = c->hm.nc;
  std::vector<double> soa((size_t)width * nc);
  for (int i = 0; i < nc; i++)
    for (int j = 0; j < width; j++)
      soa[paired ? pidx(j, nc, i) : (size_t)j * nc + i] = src[(size_t)width * i + j];
  HIPCHK(hipMemcpyAsync(dst, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));  // soa is freed on return
  return USG_OK;
}

int usg_set_state(usg_ctx *c, const double *U, const double *W, const double *turb, const double *tcons,
                  const double *mut) {
  if (!c || !U) return fail(USG_ERR_ARG, "null argument");
  const bool sst = c->prm.flow == USG_SST;
  if (sst && (!turb || !mut)) return fail(USG_ERR_ARG, "SST state needs turb and mut");
  int rc = clear_case_status(c);
  if (rc) return rc;
  if ((rc = upload_aos(c, U, 4, c->ds.U, true))) return rc;
  if (W) {
    if ((rc = upload_aos(c, W, 4, c->ds.W, true))) return rc;
  } else {
    hipLaunchKernelGGL(k_cons, dim3(c->nblk), dim3(kBlock), 0, c->stream, c->dm, c->ds, c->ds.W);
  }
  if (sst) {
    if ((rc = upload_aos(c, turb, 2, c->ds.tU, true))) return rc;
    if (tcons) {
      if ((rc = upload_aos(c, tcons, 2, c->ds.tW, true))) return rc;
    } else {
      hipLaunchKernelGGL(k_tcons, dim3(c->nblk), dim3(kBlock), 0, c->stream, c->dm, c->ds);
    }
    if ((rc = upload_aos(c, mut, 1, c->ds.mut, false))) return rc;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->have_residual = false;
  return USG_OK;
}

int usg_residual(usg_ctx *c) {
  if (!c) return fail(USG_ERR_ARG, "null ctx");
  launch_residual_any(c);
  HIPCHK(hipGetLastError());
  c->have_residual = true;
  return check_err(c);
}

static void launch_sweeps_any(usg_ctx *c, bool fuse_norm) {
  const bool mixed = c->prm.precision == USG_MIXED, batch = c->ncase > 1;
  if (mixed)
    batch ? launch_sweeps_t<float, true>(c, fuse_norm) : launch_sweeps_t<float, false>(c, fuse_norm);
  else
    batch ? launch_sweeps_t<double, true>(c, fuse_norm) : launch_sweeps_t<double, false>(c, fuse_norm);
}

static void launch_norm_ring(usg_ctx *c) {
  hipLaunchKernelGGL(k_norm_partial, dim3(c->nnb), dim3(kBlock), 0, c->stream, c->dm, c->ds.R, c->d_nb0, c->d_nb1,
                     c->d_partial);
  hipLaunchKernelGGL(k_norm_ring, dim3(c->ncase), dim3(kBlock), 0, c->stream, c->d_partial, c->d_nboff, c->d_ring,
                     kRing, c->d_ctr);
  if (c->ncase > 1) hipLaunchKernelGGL(k_add_int, dim3(1), dim3(64), 0, c->stream, c->d_ctr, 1);
}

// The norm rides on the sweep launches for a single case whose sweeps use
// kBlock-thread blocks and make at least two launches.
static bool fuse_norm(const usg_ctx *c) {
  return c->ncase == 1 && c->blk[USG_K_SWEEP] == kBlock && sweep_launches(c) >= 2;
}

// One LU-SGS iteration (LUSGSIntegrator::UpdateSerial): residual, norm into
// the ring, sweeps (the backward sweep writes the update), buffer swap.
static void launch_overset(usg_ctx *c) {
  if (c->ov_n > 0)
    hipLaunchKernelGGL(k_overset, dim3(grid(4 * c->ov_n)), dim3(kBlock), 0, c->stream, c->dm, c->ds, c->ov_n,
                       c->d_ov_recv, c->d_ov_off, c->d_ov_donor, c->d_ov_w);
}

static void one_iteration(usg_ctx *c, hipEvent_t ea, hipEvent_t eb) {
  launch_overset(c);
  launch_residual_events(c, ea, eb);
  const bool fuse = fuse_norm(c);
  if (!fuse) launch_norm_ring(c);
  launch_sweeps_any(c, fuse);
  swap_state(c);
}

static void drop_graphs(usg_ctx *c) {
  for (int i = 0; i < 2; i++) {
    if (c->gexec[i]) hipGraphExecDestroy(c->gexec[i]);
    c->gexec[i] = nullptr;
    c->gstart[i] = nullptr;
  }
}

// Captures two iterations (a -> b -> a) into an executable graph of kernel
// launches only (timing mode runs eagerly: events recorded inside a graph
// are not readable).  Capture executes nothing, so the state pointers end
// where they started.
static bool build_graph(usg_ctx *c, int slot) {
  double *start = stbuf(c);
  if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
  one_iteration(c, nullptr, nullptr);
  one_iteration(c, nullptr, nullptr);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess || !g) {
    (void)hipGetLastError();
    return false;
  }
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  c->gexec[slot] = ex;
  c->gstart[slot] = start;
  c->gstream[slot] = c->stream;
  return stbuf(c) == start;
}

// Both graphs (one per state buffer the pair starts from) are built together,
// the second with the state pointers swapped on the host only, so neither is
// ever instantiated inside a caller's timed loop after the first call.
static bool ensure_graphs(usg_ctx *c) {
  if (c->gexec[0] && c->gexec[1] && c->gstream[0] == c->stream && c->gstream[1] == c->stream) return true;
  drop_graphs(c);
  (void)fused_ready(c);  // the fused launch's tiles, before capture
  if (!build_graph(c, 0)) return false;
  swap_state(c);
  const bool ok = build_graph(c, 1);
  swap_state(c);
  return ok;
}

int usg_lusgs_step(usg_ctx *c, int n, double *res_hist) {
  if (!c || n < 0) return fail(USG_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c->stream, c->d_ctr, 0);
  int done = 0;
  while (done < n) {
    const int chunk = std::min(n - done, kRing);
    int i = 0;
    while (i < chunk) {
      if (c->use_graph && !c->timing && chunk - i >= 2) {
        if (!ensure_graphs(c)) {
          c->use_graph = false;
          drop_graphs(c);
          continue;
        }
        const int slot = c->gstart[0] == stbuf(c) ? 0 : 1;
        HIPCHK(hipGraphLaunch(c->gexec[slot], c->stream));
        i += 2;
      } else {
        std::pair<hipEvent_t, hipEvent_t> e{nullptr, nullptr};
        if (c->timing) e = next_events(c);
        one_iteration(c, e.first, e.second);
        i += 1;
      }
    }
    HIPCHK(hipGetLastError());
    if (res_hist)
      HIPCHK(hipMemcpyAsync(res_hist + (size_t)done * c->ncase, c->d_ring, (size_t)chunk * c->ncase * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c->stream, c->d_ctr, 0);
    done += chunk;
  }
  c->have_residual = false;
  return check_err(c);
}

int usg_rk5_step(usg_ctx *c, int n, double *res_hist) {
  if (!c || n < 0) return fail(USG_ERR_ARG, "bad argument");
  if (c->prm.flow == USG_SST) return fail(USG_ERR_ARG, "RK5 is provided for Euler and laminar flow");
  if (c->dm.blank || c->dt_phys > 0.0)
    return fail(USG_ERR_ARG, "RK5 is provided without overset coupling and dual time stepping");
  if (res_hist) {
    int rc = ensure_norms(c, n * c->ncase);
    if (rc) return rc;
  }
  const double alpha[5] = {1.0 / 4, 1.0 / 6, 3.0 / 8, 1.0 / 2, 1.0};
  for (int it = 0; it < n; it++) {
    hipLaunchKernelGGL(k_cons, dim3(c->nblk), dim3(kBlock), 0, c->stream, c->dm, c->ds, c->ds.W0);
    for (int st = 0; st < 5; st++) {
      launch_residual_any(c);
      if (res_hist && st == 0) launch_norm(c, c->d_norms + (size_t)it * c->ncase);
      hipLaunchKernelGGL(k_rk_stage, dim3(c->nblk), dim3(kBlock), 0, c->stream, c->dm, c->ds, c->prm.cfl_rk,
                         alpha[st], st, 0);
    }
  }
  HIPCHK(hipGetLastError());
  if (res_hist && n > 0)
    HIPCHK(hipMemcpyAsync(res_hist, c->d_norms, (size_t)n * c->ncase * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  c->have_residual = false;
  return check_err(c);
}

static int soa_to_aos(usg_ctx *c, const float *dev, int nv, double *out, long count, bool has) {
  const int nc = c->hm.nc;
  if (count != (long)nv * nc) return fail(USG_ERR_ARG, "count mismatch");
  if (!has || !dev) return fail(USG_ERR_ARG, "field not available for this flow type");
  std::vector<float> tmp((size_t)nv * nc);
  HIPCHK(hipMemcpyAsync(tmp.data(), dev, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < nc; i++)
    for (int j = 0; j < nv; j++) out[(size_t)i * nv + j] = (double)tmp[(size_t)j * nc + i];
  return USG_OK;
}

static int soa_to_aos(usg_ctx *c, const double *dev, int nv, double *out, long count, bool has, int nitems = -1) {
  const int nc = nitems >= 0 ? nitems : c->hm.nc;
  if (count != (long)nv * nc) return fail(USG_ERR_ARG, "count mismatch");
  if (!has || !dev) return fail(USG_ERR_ARG, "field not available for this flow type");
  std::vector<double> tmp((size_t)nv * nc);
  HIPCHK(hipMemcpyAsync(tmp.data(), dev, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < nc; i++)
    for (int j = 0; j < nv; j++) out[(size_t)i * nv + j] = tmp[(size_t)j * nc + i];
  return USG_OK;
}

int usg_get_field(usg_ctx *c, int field, double *out, long count) {
  if (!c || !out) return fail(USG_ERR_ARG, "null argument");
  const DevState &s = c->ds;
  const bool visc = c->prm.flow != USG_EULER, sst = c->prm.flow == USG_SST;
  switch (field) {
    case USG_F_PRIM: return pairs_to_aos(c, s.U, 4, out, count, true);
    case USG_F_CONS: return pairs_to_aos(c, s.W, 4, out, count, true);
    case USG_F_RES: return pairs_to_aos(c, s.R, 4, out, count, true);
    case USG_F_GRAD:
      return s.gradf ? pairs_to_aos(c, s.gradf, 8, out, count, true) : pairs_to_aos(c, s.grad, 8, out, count, true);
    case USG_F_LAMC: return soa_to_aos(c, s.lamc, 1, out, count, true);
    case USG_F_PHI:
      return s.phif ? pairs_to_aos(c, s.phif, 4, out, count, true) : pairs_to_aos(c, s.phi, 4, out, count, true);
    case USG_F_DW1: return pairs_to_aos(c, s.dW1, 4, out, count, true);
    case USG_F_DW: return pairs_to_aos(c, s.dW, 4, out, count, true);
    case USG_F_LAMV: return soa_to_aos(c, s.lamv, 1, out, count, visc);
    case USG_F_GRADT:
      return s.gradTf ? pairs_to_aos(c, s.gradTf, 2, out, count, visc)
                      : pairs_to_aos(c, s.gradT, 2, out, count, visc);
    case USG_F_TURB: return pairs_to_aos(c, s.tU, 2, out, count, sst);
    case USG_F_TRES: return pairs_to_aos(c, s.tR, 2, out, count, sst);
    case USG_F_MUT: return soa_to_aos(c, s.mut, 1, out, count, sst);
    case USG_F_F1: return s.F1f ? soa_to_aos(c, s.F1f, 1, out, count, sst) : soa_to_aos(c, s.F1, 1, out, count, sst);
    case USG_F_DT: return soa_to_aos(c, s.dt, 1, out, count, true);
    case USG_F_TDW: return pairs_to_aos(c, s.tdW, 2, out, count, sst);
    case USG_F_TCONS: return pairs_to_aos(c, s.tW, 2, out, count, sst);
    case USG_F_W0: return pairs_to_aos(c, s.W0, 4, out, count, true);
    case USG_F_WDIST: return soa_to_aos(c, const_cast<double *>(c->dm.wdist), 1, out, count, sst);
    case USG_F_UEDGE: {
      const int nf = c->hm.nf;
      if (count != 4L * nf) return fail(USG_ERR_ARG, "count mismatch");
      if (!c->d_uedge) {
        int rc = dalloc(c, &c->d_uedge, 4 * (size_t)nf);
        if (rc) return rc;
      }
      hipLaunchKernelGGL((visc ? k_uedge<true> : k_uedge<false>), dim3((nf + kBlock - 1) / kBlock), dim3(kBlock), 0,
                         c->stream, c->dm, c->ds, c->k, c->d_uedge);
      HIPCHK(hipGetLastError());
      return soa_to_aos(c, c->d_uedge, 4, out, count, true, nf);
    }
    default: return fail(USG_ERR_ARG, "unknown field");
  }
}

int usg_res_l2(usg_ctx *c, double *out) {
  if (!c || !out) return fail(USG_ERR_ARG, "null argument");
  int rc = ensure_norms(c, c->ncase);
  if (rc) return rc;
  launch_norm(c, c->d_norms);
  HIPCHK(hipMemcpyAsync(out, c->d_norms, c->ncase * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return USG_OK;
}

int usg_forces(usg_ctx *c, double *cl, double *cd) {
  if (!c || !cl || !cd) return fail(USG_ERR_ARG, "null argument");
  const bool visc = c->prm.flow != USG_EULER;
  hipLaunchKernelGGL((visc ? k_forces<true> : k_forces<false>), dim3(c->ncase), dim3(kBlock), 0, c->stream, c->dm,
                     c->ds, c->k, c->d_wall_faces, c->d_wall_cells, c->d_woff, c->d_force);
  HIPCHK(hipGetLastError());
  std::vector<double> f(2 * (size_t)c->ncase);
  HIPCHK(hipMemcpyAsync(f.data(), c->d_force, f.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->ncase; i++) {
    double q = 0.5 * 1.4 * c->cprm[i].mach * c->cprm[i].mach;
    double ca = std::cos(c->aoa_rad[i]), sa = std::sin(c->aoa_rad[i]);
    cl[i] = (-f[2 * i] * sa + f[2 * i + 1] * ca) / q;
    cd[i] = (f[2 * i] * ca + f[2 * i + 1] * sa) / q;
  }
  return USG_OK;
}

int usg_case_status(usg_ctx *c, int *status) {
  if (!c || !status) return fail(USG_ERR_ARG, "null argument");
  if (c->ncase == 1) {
    status[0] = 0;
    return USG_OK;
  }
  HIPCHK(hipMemcpyAsync(status, c->ds.err + 2, c->ncase * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return USG_OK;
}

int usg_case_count(usg_ctx *c, int *ncases, int *cell_off) {
  if (!c || !ncases) return fail(USG_ERR_ARG, "null argument");
  *ncases = c->ncase;
  if (cell_off)
    for (int i = 0; i <= c->ncase; i++) cell_off[i] = c->hm.case_cell[i];
  return USG_OK;
}

int usg_sweep_waves(usg_ctx *c, int *nfwd, int *nbwd) {
  if (!c || !nfwd || !nbwd) return fail(USG_ERR_ARG, "null argument");
  const HostMesh &h = c->hm;
  *nfwd = (int)h.fwd_poff.size() / (kXcd + 1);
  *nbwd = (int)h.bwd_poff.size() / (kXcd + 1);
  return USG_OK;
}

int usg_get_geometry(usg_ctx *c, double *vol, double *cc, double *fn, double *fa, double *fc, double *wdist) {
  if (!c) return fail(USG_ERR_ARG, "null ctx");
  const HostMesh &h = c->hm;
  if (vol) std::memcpy(vol, h.vol.data(), h.vol.size() * sizeof(double));
  if (cc) std::memcpy(cc, h.cc.data(), h.cc.size() * sizeof(double));
  if (fn) std::memcpy(fn, h.fn.data(), h.fn.size() * sizeof(double));
  if (fa) std::memcpy(fa, h.fa.data(), h.fa.size() * sizeof(double));
  if (fc) std::memcpy(fc, h.fc.data(), h.fc.size() * sizeof(double));
  if (wdist) std::memcpy(wdist, h.wdist.data(), h.wdist.size() * sizeof(double));
  return USG_OK;
}

int usg_mesh_build(const usg_mesh_desc *mesh, double *vol, double *cc, double *fn, double *fa, double *fc,
                   double *wdist, int *nfwd, int *nbwd) {
  if (!mesh) return fail(USG_ERR_ARG, "null mesh");
  HostMesh h;
  std::string e = build_mesh(h, mesh->nn, mesh->xy, mesh->nf, mesh->fnode, mesh->flc, mesh->frc, mesh->nc);
  if (!e.empty()) return fail(USG_ERR_MESH, e);
  if (vol) std::memcpy(vol, h.vol.data(), h.vol.size() * sizeof(double));
  if (cc) std::memcpy(cc, h.cc.data(), h.cc.size() * sizeof(double));
  if (fn) std::memcpy(fn, h.fn.data(), h.fn.size() * sizeof(double));
  if (fa) std::memcpy(fa, h.fa.data(), h.fa.size() * sizeof(double));
  if (fc) std::memcpy(fc, h.fc.data(), h.fc.size() * sizeof(double));
  if (wdist) std::memcpy(wdist, h.wdist.data(), h.wdist.size() * sizeof(double));
  if (nfwd) *nfwd = (int)h.fwd_off.size() - 1;
  if (nbwd) *nbwd = (int)h.bwd_off.size() - 1;
  return USG_OK;
}

static void dfree(usg_ctx *c, void *p) {
  if (!p) return;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
  if (it != c->allocs.end()) c->allocs.erase(it);
  (void)hipFree(p);
}

static void clear_overset(usg_ctx *c) {
  dfree(c, c->d_ov_recv);
  dfree(c, c->d_ov_off);
  dfree(c, c->d_ov_donor);
  dfree(c, c->d_ov_w);
  dfree(c, c->d_blank);
  c->d_ov_recv = c->d_ov_off = c->d_ov_donor = nullptr;
  c->d_ov_w = nullptr;
  c->d_blank = nullptr;
  c->dm.blank = nullptr;
  c->ov_n = 0;
}

// Any failure leaves the context without a coupling (as blank NULL does).
int usg_set_overset(usg_ctx *c, const int *blank, int nrecv, const int *recv, const int *off, const int *donor,
                    const double *w) {
  if (!c) return fail(USG_ERR_ARG, "null ctx");
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  clear_overset(c);
  if (!blank) return USG_OK;
  if (c->prm.flow == USG_SST) return fail(USG_ERR_ARG, "overset coupling is provided for Euler and laminar flow");
  const int nc = c->hm.nc;
  if (nrecv < 0 || nrecv > nc || (nrecv && (!recv || !off || !donor || !w)))
    return fail(USG_ERR_ARG, "bad overset stencil arguments");
  std::vector<signed char> b(nc);
  for (int i = 0; i < nc; i++) {
    if (blank[i] < -1 || blank[i] > 1) return fail(USG_ERR_ARG, "cell blanking must be 1, 0 or -1");
    b[i] = (signed char)blank[i];
  }
  // receptors: distinct cells blanked -1; donors: field cells, which the
  // interpolation never writes
  std::vector<char> seen(nc, 0);
  if (nrecv && off[0] != 0) return fail(USG_ERR_ARG, "donor offsets must start at 0");
  for (int r = 0; r < nrecv; r++) {
    if (recv[r] < 0 || recv[r] >= nc || b[recv[r]] != -1 || seen[recv[r]])
      return fail(USG_ERR_ARG, "receptor " + std::to_string(r) + " is not a distinct receptor cell");
    seen[recv[r]] = 1;
    if (off[r + 1] < off[r]) return fail(USG_ERR_ARG, "donor offsets must not decrease");
    for (int i = off[r]; i < off[r + 1]; i++)
      if (donor[i] < 0 || donor[i] >= nc || b[donor[i]] != 1 || !std::isfinite(w[i]))
        return fail(USG_ERR_ARG, "receptor " + std::to_string(r) + " has a donor that is not a field cell");
  }
  int rc;
  if ((rc = upload(c, &c->d_blank, b.data(), nc))) {
    clear_overset(c);
    return rc;
  }
  if (nrecv) {
    const int nd = off[nrecv];
    if ((rc = upload(c, &c->d_ov_recv, recv, nrecv)) || (rc = upload(c, &c->d_ov_off, off, nrecv + 1)) ||
        (rc = upload(c, &c->d_ov_donor, donor, nd)) || (rc = upload(c, &c->d_ov_w, w, nd))) {
      clear_overset(c);
      return rc;
    }
  }
  c->dm.blank = c->d_blank;
  c->ov_n = nrecv;
  HIPCHK(hipStreamSynchronize(c->stream));
  return USG_OK;
}

// Backward-difference coefficients of the current physical step into every
// case's constants (graphs captured the old ones: dropped when they change).
static int set_dual_coefficients(usg_ctx *c, double c0, double c1, double c2) {
  const double v[3] = {c0, c1, c2};
  bool changed = false;
  for (int q = 0; q < 3; q++) changed |= c->k.dts[q] != v[q];
  if (!changed) return USG_OK;
  for (int q = 0; q < 3; q++) {
    c->k.dts[q] = v[q];
    for (auto &kk : c->ck) kk.dts[q] = v[q];
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  if (c->dm.kc)
    HIPCHK(hipMemcpy((void *)c->dm.kc, c->ck.data(), c->ck.size() * sizeof(Consts), hipMemcpyHostToDevice));
  return USG_OK;
}

int usg_set_cfl(usg_ctx *c, double cfl) {
  if (!c || !(cfl > 0.0) || !std::isfinite(cfl)) return fail(USG_ERR_ARG, "bad CFL number");
  if (c->k.cfl == cfl) return USG_OK;
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  c->k.cfl = cfl;
  for (auto &kk : c->ck) kk.cfl = cfl;
  for (auto &p : c->cprm) p.cfl_lusgs = cfl;
  c->prm.cfl_lusgs = cfl;
  if (c->dm.kc)
    HIPCHK(hipMemcpy((void *)c->dm.kc, c->ck.data(), c->ck.size() * sizeof(Consts), hipMemcpyHostToDevice));
  return USG_OK;
}

int usg_set_order(usg_ctx *c, int order) {
  if (!c || (order != 1 && order != 2)) return fail(USG_ERR_ARG, "order must be 1 or 2");
  if (c->k.order == order) return USG_OK;
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  c->k.order = order;
  for (auto &kk : c->ck) kk.order = order;
  if (c->dm.kc)
    HIPCHK(hipMemcpy((void *)c->dm.kc, c->ck.data(), c->ck.size() * sizeof(Consts), hipMemcpyHostToDevice));
  return USG_OK;
}

int usg_dual_time(usg_ctx *c, double dt_phys, int order) {
  if (!c) return fail(USG_ERR_ARG, "null ctx");
  if (dt_phys > 0.0 && c->prm.flow == USG_SST)
    return fail(USG_ERR_ARG, "dual time stepping is provided for Euler and laminar flow");
  if (dt_phys > 0.0 && order != 1 && order != 2) return fail(USG_ERR_ARG, "order must be 1 or 2");
  if (dt_phys > 0.0 && !c->ds.Wt1) {
    int rc;
    if ((rc = dalloc(c, &c->ds.Wt1, 4 * (size_t)c->hm.nc)) || (rc = dalloc(c, &c->ds.Wt2, 4 * (size_t)c->hm.nc)))
      return rc;
    drop_graphs(c);   // the state view the graphs captured changed
  }
  c->dt_phys = dt_phys > 0.0 ? dt_phys : 0.0;
  c->dt_order = order;
  c->dt_steps = 0;
  return set_dual_coefficients(c, 0.0, 0.0, 0.0);
}

int usg_dual_time_advance(usg_ctx *c) {
  if (!c) return fail(USG_ERR_ARG, "null ctx");
  if (!(c->dt_phys > 0.0)) return fail(USG_ERR_ARG, "dual time stepping is off (usg_dual_time)");
  const size_t bytes = 4 * (size_t)c->hm.nc * sizeof(double);
  c->dt_steps++;
  const bool bdf2 = c->dt_order == 2 && c->dt_steps > 1;
  if (bdf2) HIPCHK(hipMemcpyAsync(c->ds.Wt2, c->ds.Wt1, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->ds.Wt1, c->ds.W, bytes, hipMemcpyDeviceToDevice, c->stream));
  const double dt = c->dt_phys;
  return bdf2 ? set_dual_coefficients(c, 1.5 / dt, -2.0 / dt, 0.5 / dt)
              : set_dual_coefficients(c, 1.0 / dt, -1.0 / dt, 0.0);
}

int usg_use_fused(usg_ctx *c, int on) {
  if (!c) return fail(USG_ERR_ARG, "null ctx");
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  c->fused = on != 0;
  if (c->fused && c->prm.precision == USG_FP64 && !build_tiles(c, c->blk[USG_K_CELL])) {
    c->fused = false;
    return fail(USG_ERR_ARG, "a tile's face record does not fit the LDS of a block");
  }
  return USG_OK;
}

int usg_get_fused(usg_ctx *c, int *on) {
  if (!c || !on) return fail(USG_ERR_ARG, "bad argument");
  *on = c->fused && c->prm.precision == USG_FP64;
  return USG_OK;
}

int usg_use_graphs(usg_ctx *c, int on) {
  if (!c) return fail(USG_ERR_ARG, "null ctx");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->use_graph = on != 0;
  if (!c->use_graph) drop_graphs(c);
  else if (!ensure_graphs(c)) {
    c->use_graph = false;
    drop_graphs(c);
    return fail(USG_ERR_HIP, "hipGraph capture of the LU-SGS iteration pair failed");
  }
  return USG_OK;
}

int usg_set_block(usg_ctx *c, int group, int threads) {
  if (!c || group < 0 || group > USG_K_SWEEP) return fail(USG_ERR_ARG, "bad kernel group");
  if (threads != 64 && threads != 128 && threads != 256)
    return fail(USG_ERR_ARG, "threads per block must be 64, 128 or 256");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->blk[group] = threads;
  drop_graphs(c);
  return USG_OK;
}

int usg_get_block(usg_ctx *c, int group, int *threads) {
  if (!c || !threads || group < 0 || group > USG_K_SWEEP) return fail(USG_ERR_ARG, "bad argument");
  *threads = c->blk[group];
  return USG_OK;
}

// Mean milliseconds of one eager LU-SGS iteration with the current blocks.
static int time_iterations(usg_ctx *c, int iters, hipEvent_t a, hipEvent_t b, float *ms) {
  one_iteration(c, nullptr, nullptr);  // warm-up (code objects, caches)
  HIPCHK(hipEventRecord(a, c->stream));
  for (int i = 0; i < iters; i++) one_iteration(c, nullptr, nullptr);
  HIPCHK(hipEventRecord(b, c->stream));
  HIPCHK(hipEventSynchronize(b));
  HIPCHK(hipEventElapsedTime(ms, a, b));
  *ms /= iters;
  return check_err(c);
}

int usg_autotune(usg_ctx *c, int iters, int *best) {
  if (!c || iters <= 0) return fail(USG_ERR_ARG, "bad argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  const int nc = c->hm.nc;
  const bool sst = c->prm.flow == USG_SST;
  // the iterations below advance the solution: keep the state to restore it
  std::vector<double *> live = {c->ds.U, c->ds.W};
  std::vector<size_t> len = {4 * (size_t)nc, 4 * (size_t)nc};
  if (sst) {
    live.insert(live.end(), {c->ds.tU, c->ds.tW, c->ds.mut});
    len.insert(len.end(), {2 * (size_t)nc, 2 * (size_t)nc, (size_t)nc});
  }
  double *const startU = c->ds.U;
  std::vector<std::vector<double>> keep(live.size());
  for (size_t i = 0; i < live.size(); i++) {
    keep[i].resize(len[i]);
    HIPCHK(hipMemcpy(keep[i].data(), live[i], len[i] * sizeof(double), hipMemcpyDeviceToHost));
  }
  hipEvent_t a = nullptr, b = nullptr;
  HIPCHK(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    hipEventDestroy(a);
    return fail(USG_ERR_HIP, "hipEventCreate");
  }
  drop_graphs(c);
  int rc = USG_OK;
  const int cand[3] = {64, 128, 256};
  // the cell group also picks the launch scheme (FP64): separate face and
  // cell launches, or the fused launch whose face tile is the cell block
  const int nscheme = c->prm.precision == USG_FP64 ? 2 : 1;
  for (int g = 0; g <= USG_K_SWEEP && rc == USG_OK; g++) {
    int pick_t = c->blk[g];
    bool pick_f = c->fused;
    float pick_ms = 1e30f;
    for (int sc = 0; sc < (g == USG_K_CELL ? nscheme : 1) && rc == USG_OK; sc++) {
      if (g == USG_K_CELL) c->fused = sc == 1;
      for (int t : cand) {
        c->blk[g] = t;
        if (c->fused && !build_tiles(c, t)) continue;  // tile record beyond the LDS limit
        float ms = 0;
        rc = time_iterations(c, iters, a, b, &ms);
        if (rc) break;
        if (ms < pick_ms) {
          pick_ms = ms;
          pick_t = t;
          pick_f = c->fused;
        }
      }
    }
    c->blk[g] = pick_t;
    c->fused = pick_f;
  }
  if (rc == USG_OK && c->fused && !build_tiles(c, c->blk[USG_K_CELL])) c->fused = false;
  hipEventDestroy(a);
  hipEventDestroy(b);
  // back to the starting buffers and state
  if (c->ds.U != startU) swap_state(c);
  for (size_t i = 0; i < live.size(); i++) {
    double *dst = i == 0 ? c->ds.U : i == 1 ? c->ds.W : i == 2 ? c->ds.tU : i == 3 ? c->ds.tW : c->ds.mut;
    HIPCHK(hipMemcpy(dst, keep[i].data(), len[i] * sizeof(double), hipMemcpyHostToDevice));
  }
  HIPCHK(hipDeviceSynchronize());
  c->have_residual = false;
  if (rc) return rc;
  if (best)
    for (int g = 0; g <= USG_K_SWEEP; g++) best[g] = c->blk[g];
  return USG_OK;
}

int usg_device_arch(char *buf, int n) {
  if (!buf || n <= 0) return fail(USG_ERR_ARG, "bad argument");
  int dev = 0;
  hipDeviceProp_t p;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipGetDeviceProperties(&p, dev));
  std::snprintf(buf, n, "%s", p.gcnArchName);
  return USG_OK;
}

int usg_timing_enable(usg_ctx *c, int on) {
  if (!c) return fail(USG_ERR_ARG, "null ctx");
  c->timing = on != 0;
  c->ev_used = 0;
  return USG_OK;
}

int usg_kernel_time(usg_ctx *c, double *total_ms, long *launches) {
  if (!c || !total_ms || !launches) return fail(USG_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  double t = 0;
  for (size_t i = 0; i < c->ev_used; i++) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[i].first, c->ev[i].second));
    t += ms;
  }
  *total_ms = t;
  *launches = (long)c->ev_used;
  c->ev_used = 0;
  return USG_OK;
}

int usg_stencil_apply(const void *dev_src, void *dev_dst, int nvar, int layout, long src_pts, long dst_pts,
                      int nrecv, const int *recv, const int *nw, const int *donor, const double *w, void *st) {
  if (!dev_src || !dev_dst || nvar <= 0 || nrecv < 0 || src_pts < 0 || dst_pts < 0 ||
      (layout != USG_ROW && layout != USG_COLUMN) || (nrecv && (!recv || !nw || !donor || !w)))
    return fail(USG_ERR_ARG, "bad argument");
  if (src_pts > 2147483647L || dst_pts > 2147483647L) return fail(USG_ERR_ARG, "point count exceeds int32");
  if (nrecv == 0) return USG_OK;
  const long nt = (long)nrecv * nvar;
  hipLaunchKernelGGL(k_stencil, dim3((unsigned)((nt + kBlock - 1) / kBlock)), dim3(kBlock), 0, (hipStream_t)st,
                     (const double *)dev_src, (double *)dev_dst, nvar, layout == USG_COLUMN ? 1 : 0, src_pts, dst_pts,
                     nrecv, recv, nw, donor, w);
  HIPCHK(hipGetLastError());
  return USG_OK;
}

int usg_stencil_apply_host(const double *src, double *dst, int nvar, int layout, long src_pts, long dst_pts,
                           int nrecv, const int *recv, const int *nw, const int *donor, const double *w) {
  if (!src || !dst || nvar <= 0 || nrecv < 0 || src_pts < 0 || dst_pts < 0 ||
      (layout != USG_ROW && layout != USG_COLUMN) || (nrecv && (!recv || !nw || !donor || !w)))
    return fail(USG_ERR_ARG, "bad argument");
  if (src_pts > 2147483647L || dst_pts > 2147483647L) return fail(USG_ERR_ARG, "point count exceeds int32");
  if (nrecv == 0) return USG_OK;
  struct Buffers {
    std::vector<void *> p;
    ~Buffers() {
      for (void *q : p) (void)hipFree(q);
    }
    void *get(size_t bytes, hipError_t &e) {
      void *q = nullptr;
      e = hipMalloc(&q, std::max<size_t>(bytes, 8));
      if (e == hipSuccess) p.push_back(q);
      return q;
    }
  } b;
  hipError_t e = hipSuccess;
  const size_t ns = (size_t)src_pts * nvar, nd = (size_t)dst_pts * nvar;
  double *ds = (double *)b.get(ns * sizeof(double), e);
  if (e == hipSuccess) {
    double *dd = (double *)b.get(nd * sizeof(double), e);
    int *dr = e == hipSuccess ? (int *)b.get((size_t)nrecv * sizeof(int), e) : nullptr;
    int *dn = e == hipSuccess ? (int *)b.get((size_t)nrecv * sizeof(int), e) : nullptr;
    int *ddn = e == hipSuccess ? (int *)b.get((size_t)nrecv * 4 * sizeof(int), e) : nullptr;
    double *dw = e == hipSuccess ? (double *)b.get((size_t)nrecv * 4 * sizeof(double), e) : nullptr;
    if (e != hipSuccess) return fail(USG_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    HIPCHK(hipMemcpy(ds, src, ns * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dd, dst, nd * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dr, recv, (size_t)nrecv * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dn, nw, (size_t)nrecv * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ddn, donor, (size_t)nrecv * 4 * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dw, w, (size_t)nrecv * 4 * sizeof(double), hipMemcpyHostToDevice));
    int rc = usg_stencil_apply(ds, dd, nvar, layout, src_pts, dst_pts, nrecv, dr, dn, ddn, dw, nullptr);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(dst, dd, nd * sizeof(double), hipMemcpyDeviceToHost));
    return USG_OK;
  }
  return fail(USG_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Staging area of usg_stencil_apply_host_noalloc: the device copies of its
// host arrays live in this fixed array instead of hipMalloc'd buffers.  Each
// GPU holds its own copy at its own address, found once per device with
// hipGetSymbolAddress; host code uses only that address (the array's name is
// a host-side stand-in).  Calls that overlap on one GPU share it and race.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

static int scratch_address(void **out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(USG_ERR_ARG, "device index beyond the staging table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(USG_ERR_HIP, "hipGetSymbolAddress returned a null staging address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return USG_OK;
}

static size_t seg256(size_t bytes) { return (bytes + 255) & ~size_t(255); }

extern "C" {

int usg_stencil_apply_host_noalloc(const double *src, double *dst, int nvar, int layout, long src_pts, long dst_pts,
                                   int nrecv, const int *recv, const int *nw, const int *donor, const double *w) {
  if (!src || !dst || nvar <= 0 || nrecv < 0 || src_pts < 0 || dst_pts < 0 ||
      (layout != USG_ROW && layout != USG_COLUMN) || (nrecv && (!recv || !nw || !donor || !w)))
    return fail(USG_ERR_ARG, "bad argument");
  if (src_pts > 2147483647L || dst_pts > 2147483647L) return fail(USG_ERR_ARG, "point count exceeds int32");
  if (nrecv == 0) return USG_OK;
  const size_t ns = (size_t)src_pts * nvar, nd = (size_t)dst_pts * nvar, nr = (size_t)nrecv;
  // layout of the staging area: src | dst | recv | nw | donor | w, each 256-byte aligned
  const size_t o_dst = seg256(ns * sizeof(double));
  const size_t o_recv = o_dst + seg256(nd * sizeof(double));
  const size_t o_nw = o_recv + seg256(nr * sizeof(int));
  const size_t o_donor = o_nw + seg256(nr * sizeof(int));
  const size_t o_w = o_donor + seg256(nr * 4 * sizeof(int));
  const size_t total = o_w + nr * 4 * sizeof(double);
  if (total > kScratchBytes)
    return fail(USG_ERR_ARG, "stencil arrays need " + std::to_string(total) +
                                 " bytes, beyond the 16 MiB staging area of usg_stencil_apply_host_noalloc");
  void *base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  unsigned char *b = (unsigned char *)base;
  double *ds = (double *)b, *dd = (double *)(b + o_dst), *dw = (double *)(b + o_w);
  int *dr = (int *)(b + o_recv), *dn = (int *)(b + o_nw), *ddn = (int *)(b + o_donor);
  HIPCHK(hipMemcpy(ds, src, ns * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dd, dst, nd * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dr, recv, nr * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dn, nw, nr * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ddn, donor, nr * 4 * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dw, w, nr * 4 * sizeof(double), hipMemcpyHostToDevice));
  int rc = usg_stencil_apply(ds, dd, nvar, layout, src_pts, dst_pts, nrecv, dr, dn, ddn, dw, nullptr);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(nullptr));
  HIPCHK(hipMemcpy(dst, dd, nd * sizeof(double), hipMemcpyDeviceToHost));
  return USG_OK;
}

int usg_copy_kernel(void *dst, const void *src, long bytes, void *st) {
  if (!dst || !src || bytes < 0 || bytes % 8) return fail(USG_ERR_ARG, "bad argument");
  long n = bytes / 8;
  int blocks = (int)std::min<long>((n + kBlock - 1) / kBlock, 16384);
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(kBlock), 0, (hipStream_t)st, (double *)dst, (const double *)src, n);
  HIPCHK(hipGetLastError());
  return USG_OK;
}

}  // extern "C"
