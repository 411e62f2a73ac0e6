// Device-side data views and the per-face arithmetic shared by the gfx950
// kernels.  Every expression keeps the reference's operand order so that,
// built with -ffp-contract=off, FP64 results reproduce the serial CPU code
// (paths relative to the reference checkout).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace usg {

constexpr int kNoSlot = -2147483647 - 1;  // empty face slot of a triangle
constexpr int kXcd = 8;                    // XCDs (and spatial parts) of an MI355X

// A work list split into kXcd parts: items[off[p] .. off[p+1]) belong to part
// p.  Block b of a launch serves part b % kXcd (workgroups are dispatched to
// the XCDs round-robin), chunk b / kXcd.  A null list is the identity (item
// i is i): the launch then saves the dependent load of its item.
struct XSlice {
  int off[kXcd + 1];
};
// (block b of the list's own blocks: b % kXcd is the part, b / kXcd the chunk)
__device__ __forceinline__ bool xcd_item_at(const int *__restrict__ items, const XSlice &sl, int b, int &item) {
  const int part = b % kXcd;
  const int i = sl.off[part] + (b / kXcd) * (int)blockDim.x + (int)threadIdx.x;
  if (i >= sl.off[part + 1]) return false;
  item = items ? items[i] : i;
  return true;
}
__device__ __forceinline__ bool xcd_item(const int *__restrict__ items, const XSlice &sl, int &item) {
  return xcd_item_at(items, sl, (int)blockIdx.x, item);
}
constexpr int kWallD = -2;
constexpr int kFarD = -3;

// Mesh, structure-of-arrays.  Cell c's face slot s lives at [s*nc + c]:
//   cf  = face id f if c is the face's left cell, -(f+1) if its right cell,
//         kNoSlot for an unused slot (triangles);
//   cnb = neighbour cell across that face, or kWallD / kFarD.
// Slots are in ascending face order (the reference's per-cell edge list).
struct DevMesh {
  int nc, nf;
  const double *vol, *ccx, *ccy, *wdist;
  const int *cf, *cnb;
  const int *flc, *frc;  // face left cell, right cell or kWallD / kFarD
  const double *fnx, *fny, *fA;
  // face offsets fdx[6][nf] = (x_f - x_L, y_f - y_L, x_f - x_R, y_f - y_R,
  // x_R - x_L, y_R - y_L) (the last four 0 on a boundary face), differenced
  // on the host with the same FP64 subtractions the kernels would make, so a
  // face reads them with its own record instead of gathering cell centres
  const double *fdx;
  // |x_R - x_L| of interior faces (viscous flows; the r_A distance of
  // SolveViscousFlux, TopologyHolderStrategy.cpp:143-188)
  const double *fdist;
  // the face normal and length again, per cell slot ([s*nc + c], the slot
  // order of cf), so that a sweep wave (a contiguous cell range) and the
  // gradient kernel read them coalesced; fslot[f] / fslot[nf + f]: the slot
  // of face f in its left / right cell's list
  const double *snx, *sny, *sA;
  const int *fslot;
  // x_f - x_c, y_f - y_c of slot s's face from cell c's centre ([s*nc + c],
  // FP64, the limiter's offsets in the gradient kernel)
  const double *sdx, *sdy;
  // FP32 copies for the mixed-precision kernels (null in FP64 mode):
  // normal, length and the offsets fdx, each rounded once from FP64 so that
  // FP32 keeps their relative precision near the wall
  const float *fnxf, *fnyf, *fAf, *fdxf;
  // batch of cases: constants of case i at kc[i], case of cell c at
  // ccase[c]; both null for a single case (the kernel argument holds them)
  const struct Consts *kc;
  const int *ccase;
  // bit 0: the cell has no lower neighbour (forward wave 0), bit 1: no
  // upper neighbour (backward wave 0); those waves are fused into others
  const unsigned char *wflag;
  // overset cell blanking (null: every cell is a field cell): 1 field,
  // 0 hole, -1 receptor.  Only field cells are updated by the sweeps and
  // enter the residual norm; receptors are set from their donors.
  const signed char *blank;
  // boundary faces: bface[nbf] (ascending), and each face's index in that
  // list (bidx[f], -1 for an interior face)
  const int *bface, *bidx;
  int nbf;
};

__device__ __forceinline__ bool frozen_cell(const DevMesh &m, int c) { return m.blank && m.blank[c] != 1; }

// State, structure-of-arrays ([var*nc + cell]).
struct DevState {
  double *U, *W;        // current primitive / conservative, [2][nc][2] (pidx)
  double *Un, *Wn;      // next iterate (double-buffered update), [2][nc][2]
  double *grad;         // [4][nc][2]: (d(var)/dx, d(var)/dy) of var v in pair v (pidx(2v+d))
  double *gradT;        // [1][nc][2]
  double *lamc, *lamv;  // [nc]
  double *mu;           // [nc] laminar viscosity at the cell
  double *phi;          // [2][nc][2] (pidx(var))
  double *R;            // [2][nc][2] (pidx)
  double *dW1, *dW;     // [2][nc][2] (pidx)
  double *dt;           // [nc] LU-SGS local time step
  double *dt_rk;        // [nc] RK5 time step, frozen at stage 0
  double *W0;           // [2][nc][2] RK5 stage origin (pidx)
  double *FF;           // face terms (q: flux 0-3, spectral 4, viscous flux 5-8, viscous
                        // spectral 9, SST 10-11), oriented out of the left cell, in
                        // pairs [pair][nf][2] (rec_pair / rec_comp in the solver)
  double *D;            // [nc] LU-SGS diagonal
  double *tD;           // [1][nc][2] SST diagonal (pidx)
  double *rA;           // [4][nc] LU-SGS r_A of each interior slot of a cell (from the
                        // neighbour across it), written by the face kernel
  // SST
  double *tU, *tW, *tUn, *tWn;  // [1][nc][2]: (k, omega) side by side (pidx)
  double *tgrad;                // [2][nc][2]: (d/dx, d/dy) of k, omega (pidx(2q+d))
  double *F1, *F2, *mut, *mutn; // [nc]
  double *tR;                   // [1][nc][2] (pidx)
  double *tdW1, *tdW;           // [1][nc][2] (pidx)
  int *err;                     // [2] invariant violation counters
  // dual time stepping: the conservative state at the two previous
  // physical time levels, W^n and W^(n-1) ([2][nc][2]; null when off)
  double *Wt1, *Wt2;
  // face states of the boundary faces for the gradient kernel, [4][nbf]
  // (k_bstate at the start of every residual)
  double *Ub;
  // FP32 storage of the mixed-precision mode (null in FP64 mode, where the
  // double arrays above hold these quantities)
  float *gradf, *phif, *gradTf, *tgradf, *F1f, *F2f, *FFf;
};

// The array holding a quantity in compute type T.
template <typename T>
__host__ __device__ __forceinline__ T *pick(double *d, float *f) {
  if constexpr (std::is_same<T, double>::value)
    return d;
  else
    return f;
}
template <typename T>
__host__ __device__ __forceinline__ const T *pickc(const double *d, const float *f) {
  if constexpr (std::is_same<T, double>::value)
    return d;
  else
    return f;
}

// Gradient-like cell arrays (grad, phi, gradT, tgrad and their FP32
// copies) keep their components in pairs, [k/2][nc][2]: component k of cell
// c at ((k/2)*nc + c)*2 + k%2.  A gather of a pair -- d/dx and d/dy of one
// variable, or two limiter values -- is then one 16-byte access.
template <typename T>
using Pair = typename std::conditional<std::is_same<T, double>::value, double2, float2>::type;
template <typename T>
__device__ __forceinline__ Pair<T> ld2(const T *a, int p, int nc, int c) {
  return reinterpret_cast<const Pair<T> *>(a)[(size_t)p * nc + c];
}
template <typename T>
__device__ __forceinline__ void st2(T *a, int p, int nc, int c, T x, T y) {
  Pair<T> v;
  v.x = x;
  v.y = y;
  reinterpret_cast<Pair<T> *>(a)[(size_t)p * nc + c] = v;
}
__host__ __device__ __forceinline__ size_t pidx(int k, int nc, int c) {
  return ((size_t)(k >> 1) * nc + c) * 2 + (k & 1);
}
// The state arrays that neighbours gather -- U, W and their second buffers,
// the sweeps' increments dW1, dW, the SST state tU, tW, tdW1, tdW, and W0,
// Wt1, Wt2 -- use the same pairs: (rho, p), (u, v) of U; (rho, rhoE),
// (rhou, rhov) of W; (k, omega) of tU.  So do the residual R, tR and the SST
// diagonal tD, which the cell kernel writes and the forward sweep reads
// together.
__device__ __forceinline__ void load4p(const double *a, int nc, int c, double *v) {
  const double2 x = ld2(a, 0, nc, c), y = ld2(a, 1, nc, c);
  v[0] = x.x;
  v[1] = x.y;
  v[2] = y.x;
  v[3] = y.y;
}
__device__ __forceinline__ void store4p(double *a, int nc, int c, const double *v) {
  st2(a, 0, nc, c, v[0], v[1]);
  st2(a, 1, nc, c, v[2], v[3]);
}

struct Consts {
  double fs[4];        // fs_primVar (rho, p, u, v) non-dimensional
  double fs_c;         // fsnd_soundSpeed = sqrt(gamma)
  double mach, re;     // fs_mach, fs_Re
  double cc;           // 110.5 / fs_Temperature (Sutherland)
  double pr, prt;      // 0.72, 0.9
  double cfl;          // pseudo-time CFL
  double fsnd_turb[2]; // free-stream (k, omega)
  double fsnd_mut;     // free-stream eddy viscosity
  double sqrt_gamma;   // sqrt(1.4)
  // dual time stepping: backward-difference coefficients over the physical
  // step, the residual gains V (dts[0] W + dts[1] W^n + dts[2] W^(n-1)) and
  // the diagonal dts[0] V; all zero when off
  double dts[3];
  // spatial order of the face reconstruction: 2 (the reference's limited
  // linear reconstruction) or 1 (the cell values, phi = 0)
  int order;
};

// Constants of the case cell c belongs to: the kernel argument for a single
// case (B = false), the batch table otherwise.
template <bool B>
__device__ __forceinline__ Consts case_consts(const DevMesh &m, const Consts &k, int c) {
  if constexpr (B)
    return m.kc[m.ccase[c]];
  else
    return k;
}

#define USG_GAMMA 1.4

template <typename T>
__device__ __forceinline__ T dot2(T a0, T a1, T b0, T b1) {
  T r = 0;
  r += a0 * b0;
  r += a1 * b1;
  return r;
}
template <typename T>
__device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }
template <typename T>
__device__ __forceinline__ T smin(T a, T b) { return (b < a) ? b : a; }

// Far-field Riemann-invariant state, BoundaryStrategy.h:262-328 (inviscid) and
// :120-187 (viscous; identical arithmetic).
__device__ __forceinline__ void farfield_state(const Consts &k, const double *Ul, double n0, double n1,
                                               double *Ue) {
  const double Gamma = USG_GAMMA;
  const double fsnd_density = k.fs[0], fsnd_pressure = k.fs[1];
  const double fsu = k.fs[2], fsv = k.fs[3];
  const double fsnd_soundSpeed = k.fs_c;
  double Vnl = dot2(Ul[2], Ul[3], n0, n1);
  double tl0 = Ul[2] - n0 * Vnl, tl1 = Ul[3] - n1 * Vnl;
  double Vnr = dot2(fsu, fsv, n0, n1);
  double tr0 = fsu - n0 * Vnr, tr1 = fsv - n1 * Vnr;
  double rho_left = Ul[0], p_left = Ul[1];
  if (fabs(Vnr) < fsnd_soundSpeed) {
    double Vnb = (Vnl + Vnr + 2 * (sqrt(Gamma * p_left / rho_left) - fsnd_soundSpeed) / (Gamma - 1)) / 2;
    double Cb = (Vnl - Vnr + 2 * (fsnd_soundSpeed + sqrt(Gamma * p_left / rho_left)) / (Gamma - 1)) *
                (Gamma - 1) / 4;
    double Sb, e0, e1;
    if (Vnr >= 0) {
      Sb = (p_left / pow(rho_left, Gamma));
      e0 = tl0 + n0 * Vnb;
      e1 = tl1 + n1 * Vnb;
    } else {
      Sb = (fsnd_pressure / pow(fsnd_density, Gamma));
      e0 = tr0 + n0 * Vnb;
      e1 = tr1 + n1 * Vnb;
    }
    Ue[0] = pow((Cb * Cb / (Gamma * Sb)), (1.0 / (Gamma - 1)));
    Ue[1] = Ue[0] * Cb * Cb / Gamma;
    Ue[2] = e0;
    Ue[3] = e1;
  } else if (Vnr >= 0) {
    Ue[0] = Ul[0]; Ue[1] = Ul[1]; Ue[2] = Ul[2]; Ue[3] = Ul[3];
  } else {
    Ue[0] = fsnd_density; Ue[1] = fsnd_pressure; Ue[2] = fsu; Ue[3] = fsv;
  }
}

// Face state U_edge of BoundaryStrategy::HandleInvicidBoundary /
// HandleViscousBoundary for a boundary face of cell state Ul.
// Returns false when the slip-wall check |t.n| > 1e-7 (:250-256) fails.
template <bool VISC>
__device__ __forceinline__ bool boundary_state(const Consts &k, int bc, const double *Ul, double n0,
                                               double n1, double *Ue) {
  if (bc == kWallD) {
    Ue[0] = Ul[0];
    Ue[1] = Ul[1];
    if (VISC) {
      Ue[2] = 0;
      Ue[3] = 0;
      return true;
    }
    double Vnl = dot2(Ul[2], Ul[3], n0, n1);
    double t0 = Ul[2] - n0 * Vnl, t1 = Ul[3] - n1 * Vnl;
    Ue[2] = t0;
    Ue[3] = t1;
    return !(dot2(t0, t1, n0, n1) > 1.0e-7);
  }
  farfield_state(k, Ul, n0, n1, Ue);
  return true;
}

// Sutherland viscosity of LaminarFluxStrategy::computeMu (LaminarFluxStrategy.h:76-96)
template <typename T>
__device__ __forceinline__ T sutherland(const Consts &k, T Tc) {
  return (T(1.0) + T(k.cc)) / (Tc + T(k.cc)) * pow(Tc, T(1.5)) * T(k.sqrt_gamma) * T(k.mach) / T(k.re);
}

// Venkatakrishnan limiter value for one face (Vankatakrishnan_Limiter.hpp:128-139)
template <typename T>
__device__ __forceinline__ T venkat(T d1max, T d1min, T delta2, T eps2) {
  // one quotient, with the bound on the side delta2 points to (lanes that
  // take different sides do not run both divisions)
  const T d = delta2 > 0 ? d1max : d1min;
  const T r = (d * d + eps2 + 2 * delta2 * d) / (d * d + 2 * delta2 * delta2 + d * delta2 + eps2);
  return (delta2 > 0 || delta2 < 0) ? r : T(1);
}

// The p* invariant of the interior HLLC flux (HLLCFluxStrategy.cpp:77-86):
// |p*_L - p*_R| < 1e-10 in FP64.  Single precision cannot meet an absolute
// 1e-10, so the FP32 flux of the mixed-precision mode checks relative to p*.
__device__ __forceinline__ bool pstar_ok(double a, double b) { return fabs(a - b) < 1e-10; }
__device__ __forceinline__ bool pstar_ok(float a, float b) {
  return fabsf(a - b) < 1.0e-4f * (fabsf(a) + fabsf(b));
}

// Interior HLLC flux (HLLCFluxStrategy.cpp:155-248).  Returns the p* check.
template <typename T>
__device__ __forceinline__ bool hllc(const T *UL, const T *UR, T n0, T n1, T area, T *F) {
  const T Gamma = T(USG_GAMMA);
  T density_left = UL[0], pressure_left = UL[1];
  T vl0 = UL[2], vl1 = UL[3];
  T Vn_left = dot2(vl0, vl1, n0, n1);
  T soundspeed_left = sqrt(Gamma * pressure_left / density_left);
  T vl2 = dot2(vl0, vl1, vl0, vl1);
  T enthalpy_left = pressure_left * Gamma / ((Gamma - 1) * density_left) + T(0.5) * vl2;
  T density_righ = UR[0], pressure_righ = UR[1];
  T vr0 = UR[2], vr1 = UR[3];
  T Vn_righ = dot2(vr0, vr1, n0, n1);
  T soundspeed_righ = sqrt(Gamma * pressure_righ / density_righ);
  T vr2 = dot2(vr0, vr1, vr0, vr1);
  T enthalpy_righ = pressure_righ * Gamma / ((Gamma - 1) * density_righ) + T(0.5) * vr2;
  T efficient_left = sqrt(density_left) / (sqrt(density_left) + sqrt(density_righ));
  T efficient_righ = 1 - efficient_left;
  T enthalpy_tilda = efficient_left * enthalpy_left + efficient_righ * enthalpy_righ;
  T Vn_tilda = efficient_left * Vn_left + efficient_righ * Vn_righ;
  // the reference's pow(Vn, 2) as the one rounded product: the correctly
  // rounded square, where a general pow costs ~200 instructions here
  T soundspeed_tilda = sqrt((Gamma - 1) * (enthalpy_tilda - T(0.5) * (Vn_tilda * Vn_tilda)));
  T Sl = smin(Vn_left - soundspeed_left, Vn_tilda - soundspeed_tilda);
  T Sr = smax(Vn_tilda + soundspeed_tilda, Vn_righ + soundspeed_righ);
  T Sm = (density_righ * Vn_righ * (Sr - Vn_righ) - density_left * Vn_left * (Sl - Vn_left) + pressure_left -
               pressure_righ) /
              (density_righ * (Sr - Vn_righ) - density_left * (Sl - Vn_left));
  T p_star = density_left * (Vn_left - Sl) * (Vn_left - Sm) + pressure_left;
  T p_star_scratch = density_righ * (Vn_righ - Sr) * (Vn_righ - Sm) + pressure_righ;
  bool ok = pstar_ok(p_star_scratch, p_star);
  if (Sl > T(0.0)) {
    T rhoE = pressure_left / (Gamma - 1) + T(0.5) * density_left * vl2;
    F[0] = Vn_left * density_left * area;
    F[1] = Vn_left * (rhoE + pressure_left) * area;
    F[2] = (Vn_left * density_left * vl0 + pressure_left * n0) * area;
    F[3] = (Vn_left * density_left * vl1 + pressure_left * n1) * area;
  } else if (Sr < T(0.0)) {
    T rhoE = pressure_righ / (Gamma - 1) + T(0.5) * density_righ * vr2;
    F[0] = Vn_righ * density_righ * area;
    F[1] = Vn_righ * (rhoE + pressure_righ) * area;
    F[2] = (Vn_righ * density_righ * vr0 + pressure_righ * n0) * area;
    F[3] = (Vn_righ * density_righ * vr1 + pressure_righ * n1) * area;
  } else {
    T SRef, Vn_ref, density_ref, pressure_ref, v0, v1, vv;
    if (Sm < T(0.0)) {
      SRef = Sr; Vn_ref = Vn_righ; density_ref = density_righ; pressure_ref = pressure_righ;
      v0 = vr0; v1 = vr1; vv = vr2;
    } else {
      SRef = Sl; Vn_ref = Vn_left; density_ref = density_left; pressure_ref = pressure_left;
      v0 = vl0; v1 = vl1; vv = vl2;
    }
    T rhoE_ref = pressure_ref / (Gamma - 1) + T(0.5) * density_ref * vv;
    T f0 = (SRef - Vn_ref) * density_ref / (SRef - Sm);
    T f1 = ((SRef - Vn_ref) * rhoE_ref - pressure_ref * Vn_ref + p_star * Sm) / (SRef - Sm);
    T f2 = ((SRef - Vn_ref) * density_ref * v0 + (p_star - pressure_ref) * n0) / (SRef - Sm);
    T f3 = ((SRef - Vn_ref) * density_ref * v1 + (p_star - pressure_ref) * n1) / (SRef - Sm);
    F[0] = f0 * Sm * area;
    F[1] = (f1 + p_star) * Sm * area;
    F[2] = (f2 * Sm + p_star * n0) * area;
    F[3] = (f3 * Sm + p_star * n1) * area;
  }
  return ok;
}

// Wall / far-field flux of HLLCFluxStrategy::computeFlux (HLLCFluxStrategy.cpp:126-154)
__device__ __forceinline__ void boundary_flux(int bc, const double *Ue, double n0, double n1, double area,
                                              double *F) {
  const double Gamma = USG_GAMMA;
  if (bc == kWallD) {
    F[0] = 0 * area;
    F[1] = 0 * area;
    double P = Ue[1];
    F[2] = P * n0 * area;
    F[3] = P * n1 * area;
  } else {
    double D = Ue[0], P = Ue[1], v0 = Ue[2], v1 = Ue[3];
    double rhoE = P / (Gamma - 1) + 0.5 * D * dot2(v0, v1, v0, v1);
    double Vn = dot2(v0, v1, n0, n1);
    F[0] = Vn * D * area;
    F[1] = Vn * (rhoE + P) * area;
    F[2] = (Vn * D * v0 + P * n0) * area;
    F[3] = (Vn * D * v1 + P * n1) * area;
  }
}

// ConservativeParam2ConvectiveFlux (TopologyHolderStrategy.cpp:129-141)
__device__ __forceinline__ void cons_flux(const double *W, double n0, double n1, double *Fc) {
  const double Gamma = USG_GAMMA;
  double u = W[2] / W[0], v = W[3] / W[0];
  double Vn = dot2(u, v, n0, n1);
  double P = (Gamma - 1) * (W[1] - 0.5 * W[0] * dot2(u, v, u, v));
  Fc[0] = W[0] * Vn;
  Fc[1] = (W[1] + P) * Vn;
  Fc[2] = W[2] * Vn + P * n0;
  Fc[3] = W[3] * Vn + P * n1;
}

// UpdatePrimitiveVariables (TopologyHolderStrategy.cpp:89-109)
__device__ __forceinline__ void prim_of(const double *W, double *U) {
  const double Gamma = USG_GAMMA;
  U[0] = W[0];
  U[2] = W[2] / U[0];
  U[3] = W[3] / U[0];
  double VM = 0;
  VM += U[2] * U[2];
  VM += U[3] * U[3];
  U[1] = (Gamma - 1) * (W[1] - 0.5 * U[0] * VM);
}

// initializeConsData (TopologyHolderStrategy.cpp:48-66)
__device__ __forceinline__ void cons_of(const double *U, double *W) {
  const double Gamma = USG_GAMMA;
  double VM = 0;
  VM += U[2] * U[2];
  W[2] = U[0] * U[2];
  VM += U[3] * U[3];
  W[3] = U[0] * U[3];
  W[0] = U[0];
  W[1] = U[1] / (Gamma - 1) + 0.5 * U[0] * VM;
}

}  // namespace usg
