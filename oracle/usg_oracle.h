/*
 * usg_oracle.h -- CPU FP64 oracle for the UnstructSolver residual-and-update
 * path.  TEST INFRASTRUCTURE ONLY: this library is the checker that tests/,
 * the smoke entry and the CPU-baseline leg of bench.py compare against.  The
 * product (tmp_amd, libusg) never links, loads or calls it.
 *
 * Every function restates, loop for loop and expression for expression, the
 * serial CPU code of the reference (SenhuWong/tmp, src/UnstructSolver/),
 * keeping the reference's array-of-structures layout:
 *   U, W, R  : [cell*NE + var]  with U = (rho, p, u, v), W/R = (rho, rhoE, rhou, rhov)
 *   gradU    : [(cell*NE + var)*2 + dim]
 *   U_edge   : [face*NE + var]
 * Compile with -O2 -ffp-contract=off (no FMA contraction, as the reference's
 * x86-64 build without -march flags).
 */
#ifndef USG_ORACLE_H
#define USG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_WALL = -2, ORC_FARFIELD = -3 };
enum { ORC_EULER = 0, ORC_LAMINAR = 1, ORC_SST = 2 };

typedef struct orc_mesh {
  int nn, nf, nc;
  /* inputs (copied) */
  double *xy;        /* [nn*2] node coordinates                        */
  int *fnode;        /* [nf*2] face nodes, 0-based, file order          */
  int *flc, *frc;    /* [nf] left cell / right cell or ORC_WALL/FARFIELD */
  /* derived geometry (edge3d/cell3d restatement) */
  int *cnode;        /* [nc*4] cell vertices (Brick4Cobalt2D order)     */
  int *cnv;          /* [nc]  vertex count                               */
  int *cface;        /* [nc*4] cell faces, ascending                     */
  int *cnf;          /* [nc]  face count                                 */
  double *vol;       /* [nc]                                             */
  double *cc;        /* [nc*2] vertex-mean centre                        */
  double *fn;        /* [nf*2] unit normal                               */
  double *fa;        /* [nf]   face length                               */
  double *fc;        /* [nf*2] face midpoint                             */
  double *wdist;     /* [nc]   projected wall distance                   */
  double *wdist_wall; /* [nc]  distance to the nearest wall face segment (SST) */
} orc_mesh;

typedef struct orc_solver {
  orc_mesh *m;
  int flow;          /* ORC_EULER / ORC_LAMINAR / ORC_SST */
  int neq;           /* 4 */
  /* freestream (TopologyHolderStrategy members) */
  double R_gas, Gamma, PI;
  double fs_mach, fs_AOA, fs_density, fs_pressure, fs_eigen_Length;
  double fs_soundSpeed, fs_velocity_magnitude, fs_velocity_components[3];
  double fs_Temperature, fs_Pr, fs_Prt, fs_mu, fs_Re;
  double fsnd_density, fsnd_pressure, fsnd_soundSpeed, fsnd_velocity_components[3];
  double fs_primVar[5];
  /* flow arrays */
  double *U, *U_edge, *gradU, *gradT, *R, *lamc, *lamv, *dt;
  double *Umax, *Umin, *phi, *UL, *UR;
  double *mu, *mu_edge;
  /* LU-SGS */
  double cfl_lusgs;
  double *D, *W, *dW1, *dW;
  /* RK5 */
  double cfl_rk, rk_alpha[5];
  double *W0, *Wnow;
  /* SST (k, omega) */
  double *tU, *tUedge, *tgrad, *F1, *F2, *mut, *mut_face, *tFc, *tFv, *tR;
  double *tD, *tW, *tdW1, *tdW;
  double fs_mut, fsnd_mut, fsnd_mu, fs_turb[2], fsnd_turb[2];
  /* invariant counters */
  long hllc_pstar_violations;
  /* viscous momentum flux of each face (tau . n A), for the wall force */
  double *Fv_face;
  /* the whole viscous face flux (0, (u.tau - q).n A, tau.n A) [nf*4], for tests */
  double *Fv_all;
  /* dual time stepping (no counterpart in the reference): physical step,
   * order, steps taken, BDF coefficients / dt, earlier time levels [nc*4] */
  double dt_phys;
  int dt_order, dt_steps;
  double dts[3];
  double *Wt1, *Wt2;
  /* overset cell blanking (NULL: every cell a field cell): 1 field, 0 hole,
   * -1 receptor; only field cells are updated and enter the norm */
  int *blank;
  /* reconstruction order: 2 (the reference's limited linear) or 1 (cell values) */
  int order;
} orc_solver;

/* mesh / geometry (edge3d_int.h, cell3d_int.h, cellVolume.cpp, CellBrick.cpp) */
orc_mesh *orc_mesh_create(int nn, const double *xy, int nf, const int *fnode,
                          const int *flc, const int *frc, int nc);
void orc_mesh_destroy(orc_mesh *m);

/* solver */
orc_solver *orc_solver_create(orc_mesh *m, int flow, double ma, double aoa_deg, double re);
void orc_solver_destroy(orc_solver *s);
void orc_init_data(orc_solver *s);               /* freestream cold start      */
void orc_set_prim(orc_solver *s, const double *U); /* warm start, AoS [nc*4] */
void orc_residual(orc_solver *s);                /* preprocessAdvanceSerial     */
void orc_lusgs_step(orc_solver *s);              /* LUSGSIntegrator::UpdateSerial */
void orc_rk5_step(orc_solver *s);                /* RK5 singleStepSerial        */
double orc_density_residual_l2(orc_solver *s);   /* ||R_rho / V||_2             */
double orc_dw_rho_sq(orc_solver *s);             /* sum dW_rho^2 (postprocessUpdate) */
void orc_forces(orc_solver *s, double *cl, double *cd);
void orc_sst_eddy_viscosity(orc_solver *s);      /* mu_t of the current state */
/* dual time stepping: BDF1 on the first physical step, then BDF2 (order 2) */
void orc_dual_time(orc_solver *s, double dt_phys, int order);
void orc_dual_time_advance(orc_solver *s);
/* overset: cell blanking (NULL clears) and receptor states U[n*4] of cells idx[n] */
void orc_set_blank(orc_solver *s, const int *blank);
void orc_set_cells(orc_solver *s, int n, const int *idx, const double *U);
/* reconstruction order (no counterpart in the reference, which is second order) */
void orc_set_order(orc_solver *s, int order);
/* standalone pieces for invariant tests */
void orc_lusgs_sweeps(orc_solver *s);
double orc_cell_volume(const double *xy, int nv);
void orc_prim_to_cons(const double *U, double *W);
void orc_cons_to_prim(const double *W, double *U);
void orc_hllc_face(const double *UL, const double *UR, const double *n, double area,
                   double *flux, int *pstar_ok);
void orc_lusgs_levels(const orc_mesh *m, int *flevel, int *blevel, int *nfl, int *nbl);

/* overset stencil: donor weights of a point in a triangle / quadrilateral
 * (xv[nvert*2]) and the interpolation q_dst[recv] = sum w q_src[donor] */
int orc_nodal_weights(int nvert, const double *xv, const double *xp, double *frac);
void orc_stencil_apply(const double *src, double *dst, int nvar, int nrecv, const int *recv, const int *nw,
                       const int *donor, const double *w);

#ifdef __cplusplus
}
#endif
#endif
