/*
 * usg.h -- C ABI of the MI355X (gfx950) UnstructSolver residual-and-update
 * path.  Plain pointers and sizes only; every call returns USG_OK (0) or a
 * negative status and never throws.  Host arrays use the reference's
 * array-of-structures layout (value of variable j of cell k at [k*nvar + j]),
 * so a TopologyHolderStrategy adapter can hand its getU()/getResidual()
 * buffers straight through; device storage is structure-of-arrays.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the reference checkout, SenhuWong/tmp).
 */
#ifndef USG_H
#define USG_H

#ifdef __cplusplus
extern "C" {
#endif

#define USG_OK 0
#define USG_ERR_ARG -1      /* bad argument (null pointer, size, enum)          */
#define USG_ERR_MESH -2     /* invalid mesh (left cell < 0, unknown BC tag, ...) */
#define USG_ERR_HIP -3      /* HIP runtime error                                */
#define USG_ERR_NUMERIC -4  /* invariant violated on the device (HLLC p*, slip wall) */
#define USG_ERR_NODEVICE -5 /* no GPU visible                                   */

/* boundary codes of a face's right cell: toolBox/edge3d_int.h:33-40 */
#define USG_WALL -2
#define USG_FARFIELD -3

/* flow type: WrapSolver::setParams "Euler2D" / "ViscousFlow2D" (SolverWrapper.h:30-41);
 * USG_SST is ViscousFlow2D with isLaminar() == false (ViscousFlow2D.h:156-159). */
#define USG_EULER 0
#define USG_LAMINAR 1
#define USG_SST 2

/* precision of the residual kernels */
#define USG_FP64 0  /* every kernel in FP64 (the drop-in)                      */
#define USG_MIXED 1 /* gradients, limiter and fluxes in FP32; residual sums,
                       LU-SGS diagonal inverse and norms in FP64.  Limit: the
                       SST k residual is a near-cancellation of production and
                       destruction; one FP32 evaluation moves it by up to 0.5
                       of its largest value on the 1M-cell airfoil (the flow
                       equations by < 1e-3; bench.py reports it).  Converged
                       there (first-order reconstruction, 6000 iterations),
                       the mixed solution's C_L/C_D stay within 4e-6 of
                       FP64's and its fields within 2.5e-4. */

/* fields readable with usg_get_field (host layout AoS, see each comment) */
#define USG_F_PRIM 0     /* U   [nc*4]  (rho, p, u, v)     Euler2D::getU          */
#define USG_F_CONS 1     /* W   [nc*4]  (rho, rhoE, rhou, rhov)  LUSGS W_scratch */
#define USG_F_RES 2      /* R   [nc*4]                      getResidual           */
#define USG_F_GRAD 3     /* gradU [nc*4*2]                  getGradientPrimitive  */
#define USG_F_LAMC 4     /* spectral radius [nc]            getSpectrumConvective */
#define USG_F_PHI 5      /* limiter [nc*4]                  Vankatakrishnan getLimiter */
#define USG_F_DW1 6      /* forward-sweep increment [nc*4]  LUSGS d_deltaW1       */
#define USG_F_DW 7       /* backward-sweep increment [nc*4] LUSGS d_deltaW        */
#define USG_F_LAMV 8     /* viscous spectral radius [nc]    getSpectrumViscous    */
#define USG_F_GRADT 9    /* temperature gradient [nc*2]     getGradientT          */
#define USG_F_TURB 10    /* (k, omega) [nc*2]               SSTkomegaModel::getU  */
#define USG_F_TRES 11    /* turbulence residual [nc*2]      SSTkomegaModel::getResidual */
#define USG_F_MUT 12     /* eddy viscosity [nc]             SSTkomegaModel vCellMut */
#define USG_F_F1 13      /* SST blending F1 [nc]                                  */
#define USG_F_DT 14      /* local time step [nc]            getStableTime         */
#define USG_F_TDW 15     /* turbulence backward increment [nc*2]                   */
#define USG_F_TCONS 16   /* (rho k, rho omega) [nc*2]       SSTkomegaModel W      */
#define USG_F_W0 17      /* W at the start of the last RK5 step [nc*4]  W_scratch */
#define USG_F_UEDGE 18   /* face states of the current U [nf*4]  getUEdge (computed on request) */
#define USG_F_WDIST 19   /* SST: distance to the nearest wall face segment [nc] (the
                            reference's projected distance is usg_get_geometry's wdist) */

typedef struct usg_mesh_desc {
  int nn;              /* nodes                                              */
  int nf;              /* faces                                              */
  int nc;              /* cells                                              */
  const double *xy;    /* [nn*2] node coordinates                            */
  const int *fnode;    /* [nf*2] 0-based face nodes; normal (dy,-dx) out of left cell */
  const int *flc;      /* [nf] left cell (0-based)                           */
  const int *frc;      /* [nf] right cell, or USG_WALL / USG_FARFIELD        */
} usg_mesh_desc;

typedef struct usg_params {
  int flow;            /* USG_EULER / USG_LAMINAR / USG_SST                  */
  int precision;       /* USG_FP64 / USG_MIXED                               */
  double mach;         /* free-stream Mach number                            */
  double aoa_deg;      /* angle of attack in degrees                         */
  double reynolds;     /* Reynolds number (viscous flows)                    */
  double cfl_lusgs;    /* 400 in the reference (LUSGS_Integrator.h:18)       */
  double cfl_rk;       /* 0.8 in the reference (RungeKuttaStrategy.h:125)    */
} usg_params;

typedef struct usg_ctx usg_ctx;

/* Context = one UnstructTopologyHolder + flow strategy + integrator on one
 * HIP stream.  Replaces WrapSolver::setTopology/setParams/init
 * (SolverWrapper.h:22-68): builds geometry (UnstructBlock.h:572-586), the
 * free stream (TopologyHolderStrategy.cpp:325-351) and the LU-SGS wave
 * schedule, uploads SoA arrays.  Contexts are independent.  A mesh (or a
 * batch's union) may hold at most 268435455 cells and 357913941 faces (the
 * kernels' 32-bit array offsets); a larger one returns USG_ERR_ARG before
 * any array is read. */
int usg_ctx_create(const usg_mesh_desc *mesh, const usg_params *params, usg_ctx **out);
/* Batched many-case context: ncases independent meshes with their own free
 * streams (mach, aoa, reynolds may differ; flow type, precision and CFL
 * numbers must agree) solved together as one block-diagonal union, cells
 * and faces case-major.  Each case keeps its own geometry, wall distance
 * and numbering, so its results equal a usg_ctx_create context's bit for
 * bit; the union's LU-SGS waves are the cases' waves side by side, so a
 * batch needs no more launches per iteration than one case.  Replaces
 * running WrapSolver (SolverWrapper.h:7-95) once per case, as Orchestra
 * does per block (Orchestra.cpp:121-153).  Everything below then works on
 * the union: fields are case-major, and the per-case outputs
 * (res_hist, usg_res_l2, usg_forces) hold one value per case. */
int usg_batch_create(const usg_mesh_desc *meshes, const usg_params *params, int ncases, usg_ctx **out);
/* Number of cases and, if cell_off is not NULL, their cell ranges
 * (cell_off[ncases+1]: case i owns cells [cell_off[i], cell_off[i+1])). */
int usg_case_count(usg_ctx *ctx, int *ncases, int *cell_off);
/* Per-case status of a batch, status[ncases]: 0 = running, bit 0 = the HLLC
 * p* check failed (HLLCFluxStrategy.cpp:77-86), bit 1 = the slip-wall check
 * failed (BoundaryStrategy.h:250-256).  Where a single context's call fails
 * with USG_ERR_NUMERIC, a batch freezes the failed case at its state of that
 * iteration and carries on with the others; usg_set_prim and
 * usg_init_freestream clear the status. */
int usg_case_status(usg_ctx *ctx, int *status);
int usg_ctx_destroy(usg_ctx *ctx);
/* hipStream_t to launch on (NULL = the context's own stream). */
int usg_set_stream(usg_ctx *ctx, void *hip_stream);
int usg_sync(usg_ctx *ctx);

/* Cold start: Euler2D::initializeData (Euler2D.cpp:153-171) and
 * LUSGSIntegrator::initializeData (LUSGS_Integrator.cpp:107-111); SST adds
 * SSTkomegaModel::InitializeturbulenceVariable (SSTKWStrategy.h:582-613). */
int usg_init_freestream(usg_ctx *ctx);
/* Warm start from primitive U [nc*4] (and, for SST, (k,omega) [nc*2], may be
 * NULL: (k, omega) stay as they are).  For SST the conservative (rho k,
 * rho omega) is recomputed from (k, omega) and the new density either way. */
int usg_set_prim(usg_ctx *ctx, const double *U, const double *turb);
/* Restart: the whole pseudo-time state, as usg_get_field reads it --
 * U [nc*4] (USG_F_PRIM), W [nc*4] (USG_F_CONS; NULL: computed from U), and
 * for SST turb [nc*2] (USG_F_TURB), tcons [nc*2] (USG_F_TCONS; NULL: from
 * turb) and mut [nc] (USG_F_MUT).  An iteration carries exactly this state
 * (W is updated and U derived from it), so a context restarted from a
 * state read after iteration n continues with the iterates of the original
 * run.  No counterpart in the reference (it has no restart). */
int usg_set_state(usg_ctx *ctx, const double *U, const double *W, const double *turb, const double *tcons,
                  const double *mut);

/* preprocessAdvanceSerial (Euler2D.h:133-140; ViscousFlow2D.h:192-212):
 * boundary states, gradients, spectral radii, limiter, HLLC (+ viscous,
 * + SST) residual of the current state. */
int usg_residual(usg_ctx *ctx);
/* n LU-SGS pseudo-time iterations: LUSGSIntegrator::singleStepSerial /
 * UpdateSerial (LUSGS_Integrator.cpp:124-129,228-255).  res_hist (may be
 * NULL) receives ||R_rho/V||_2 of each iteration's residual, per case:
 * res_hist[it * ncases + case]. */
int usg_lusgs_step(usg_ctx *ctx, int n, double *res_hist);
/* n RK5 steps: RungeKuttaStrategy::singleStepSerial (RungeKuttaStrategy.h:28-72);
 * res_hist as for usg_lusgs_step (the stage-0 residual of each step). */
int usg_rk5_step(usg_ctx *ctx, int n, double *res_hist);

/* Copy a field to the host (AoS layout of the reference); count = elements. */
int usg_get_field(usg_ctx *ctx, int field, double *out, long count);
/* ||R_rho / V||_2 of the last residual (the convergence norm), out[ncases]. */
int usg_res_l2(usg_ctx *ctx, double *out);
/* Lift and drag coefficients of the current state, cl[ncases], cd[ncases]:
 * the momentum flux out of the fluid through the wall faces -- p n A with
 * the wall cell's pressure, minus the viscous flux tau . n A for viscous
 * flow, with the wall cell's velocity gradient and the face stress computed
 * from the current state exactly as a residual evaluation would (no other
 * field of the context changes) -- rotated by the angle of attack and
 * scaled by 1/2 gamma M^2.  The reference integrates no forces; this is the
 * definition used here. */
int usg_forces(usg_ctx *ctx, double *cl, double *cd);

/* Mesh facts: number of LU-SGS waves (forward, backward) of the numbering. */
int usg_sweep_waves(usg_ctx *ctx, int *nfwd, int *nbwd);
/* Host geometry as the reference computes it (cell3d/edge3d/UnstructBlock):
 * vol[nc], cc[nc*2], fn[nf*2], fa[nf], fc[nf*2], wdist[nc]; any may be NULL. */
int usg_get_geometry(usg_ctx *ctx, double *vol, double *cc, double *fn, double *fa,
                     double *fc, double *wdist);

/* Host-only mesh preparation (no GPU needed): the same geometry and LU-SGS
 * wave schedule usg_ctx_create builds, for inspection and tests. */
int usg_mesh_build(const usg_mesh_desc *mesh, double *vol, double *cc, double *fn, double *fa,
                   double *fc, double *wdist, int *nfwd, int *nbwd);

/* Host-only colour-blocked renumbering, so that the LU-SGS sweeps (which
 * follow the cell numbering, LUSGS_Integrator.cpp:492-561) run one wave per
 * colour: 3 or 4 waves where a file numbering gives hundreds of levels.
 * Cells are coloured (three colours, a fourth for a few cells) and sorted by
 * (colour, spatial part, Hilbert key); faces get lC < rC (node order flipped
 * with them) and are sorted by (part of lC, lC, rC).  Outputs, caller
 * allocated: perm[nc] (new cell -> old cell), the renumbered faces
 * fnode[2*nf], flc[nf], frc[nf] (boundary tags kept), and the number of
 * colours.  Node coordinates are unchanged.  The same permutation as
 * tmp_amd/mesh.py prepare().  No counterpart in the reference (which sweeps
 * its file numbering); a renumbered mesh is an ordinary mesh for it. */
int usg_mesh_renumber(const usg_mesh_desc *mesh, int *perm, int *fnode, int *flc, int *frc, int *ncolours);

/* LU-SGS iterations run as replays of a captured two-iteration hipGraph
 * (default on); off = eager launches.  Results are identical either way. */
int usg_use_graphs(usg_ctx *ctx, int on);
/* FP64 residuals: on = the face fluxes and the cell gathers in one launch,
 * the face record held in LDS per tile of cells; off (default) = a face
 * launch that writes the face record to HBM and a cell launch that reads it
 * back.  Results are identical either way; which is faster depends on the
 * mesh (the fused launch computes the faces between tiles twice and holds
 * fewer waves per CU), so the autotuner picks it.  Fails with USG_ERR_ARG
 * (and stays off) when a tile's face record would not fit the LDS of a
 * block.  usg_get_fused reports the scheme in use (0 in mixed precision).
 * No counterpart in the reference (a launch choice). */
int usg_use_fused(usg_ctx *ctx, int on);
int usg_get_fused(usg_ctx *ctx, int *on);

/* Threads per block of each kernel group (64, 128 or 256; default 256):
 * the gradient/limiter kernel, the face flux kernel, the cell gather kernel
 * and the sweep waves.  No counterpart in the reference (a GPU launch
 * parameter). */
#define USG_K_GRAD 0
#define USG_K_FACE 1
#define USG_K_CELL 2
#define USG_K_SWEEP 3
int usg_set_block(usg_ctx *ctx, int group, int threads);
int usg_get_block(usg_ctx *ctx, int group, int *threads);
/* Autotuner: for each group in turn, times `iters` eager LU-SGS iterations
 * with 64, 128 and 256 threads per block and keeps the fastest; in FP64 the
 * cell group also tries the fused face-cell launch (usg_use_fused) at each
 * size, whose face tile is the cell block.  The state is restored
 * afterwards, so tuning does not change the solution.  best[4] (may be
 * NULL) receives the chosen sizes; the chosen scheme stays set
 * (usg_get_fused).  No counterpart in the reference. */
int usg_autotune(usg_ctx *ctx, int iters, int *best);
/* gcnArchName of the current device (key of autotuning caches). */
int usg_device_arch(char *buf, int n);

/* Kernel timing: when enabled, every launch of the dominant kernel (the face
 * flux kernel) is bracketed by HIP events on its stream; iterations then run
 * eagerly (not as graph replays).  The total and the count are read (and
 * reset) with usg_kernel_time. */
int usg_timing_enable(usg_ctx *ctx, int on);
int usg_kernel_time(usg_ctx *ctx, double *total_ms, long *launches);

/* Overset donor-stencil apply (MeshBlock::getInterpolatedSolution +
 * updateSolnData, UniTioga/dataUpdate.cpp:6-133) on device arrays:
 *   dst[recv[r]][k] = sum_{m < nw[r]} w[r*4+m] * src[donor[r*4+m]][k]
 * for r < nrecv, k < nvar, summed in m order from 0.  layout USG_ROW stores
 * point i's variables at q[i*nvar + k] (the reference's ROW), USG_COLUMN at
 * q[k*npts + i] (the solver's own layout); src_pts / dst_pts are the arrays'
 * point counts.  recv, nw, donor, w are device arrays.  Entries whose
 * indices fall outside the arrays are skipped.  If src and dst are the same
 * array, no receptor may also be a donor (receptors are fringe cells,
 * donors field cells, as in an overset assembly). */
#define USG_ROW 0
#define USG_COLUMN 1
int usg_stencil_apply(const void *dev_src, void *dev_dst, int nvar, int layout, long src_pts, long dst_pts,
                      int nrecv, const int *recv, const int *nw, const int *donor, const double *w,
                      void *hip_stream);
/* usg_stencil_apply on host arrays (copied to the GPU and back; synchronous):
 * the form tioga::dataUpdate (UniTioga/tioga.cpp:326-471) has, updating the
 * solution arrays registered with registerSolution in place. */
int usg_stencil_apply_host(const double *src, double *dst, int nvar, int layout, long src_pts, long dst_pts,
                           int nrecv, const int *recv, const int *nw, const int *donor, const double *w);
/* usg_stencil_apply_host without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate GPU memory (code
 * under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The device copies of the arrays are staged in a fixed 16 MiB
 * array that the library holds on each GPU it runs on until the process
 * ends; calls whose arrays need more than that (src and dst at
 * src_pts*nvar and dst_pts*nvar doubles, 24 bytes of indices and 32 bytes of
 * weights per receptor, each part rounded up to 256 bytes) return
 * USG_ERR_ARG.  It calls no hipMalloc/hipFree.  Every call on a GPU shares
 * that GPU's staging area, so calls that overlap on one GPU (from several
 * host threads) race; serialise them.  Results equal usg_stencil_apply_host's
 * bit for bit. */
int usg_stencil_apply_host_noalloc(const double *src, double *dst, int nvar, int layout, long src_pts,
                                   long dst_pts, int nrecv, const int *recv, const int *nw, const int *donor,
                                   const double *w);
/* Donor weights of n receptor points (host arrays): point r lies in a cell
 * of nvert[r] (3 or 4) vertices xv[r*8 .. r*8+2*nvert[r]) and gets
 * w[r*4 .. r*4+nvert[r]) -- computeNodalWeight2D (mathTools/math.cpp:229-312):
 * barycentric for a triangle, bilinear (Newton, |F| <= 1e-14, 500 steps) for
 * a quadrilateral.  Host only, no GPU needed. */
int usg_nodal_weights(int n, const int *nvert, const double *xv, const double *xp, double *w);

/* ---- outputs (host only, no GPU needed) ---------------------------------
 * U below is a host array of primitive states U[nc*4] = (rho, p, u, v) per
 * cell, as usg_get_field(USG_F_PRIM) returns it; fs[4] the non-dimensional
 * free stream (rho, p, u, v). */
/* Cell vertices in the reference's order (cell3d::pointInd after
 * Brick4Cobalt2D::insert, CellBrick.cpp:13-136): cnode[nc*4] (-1 padded),
 * cnv[nc] in {3, 4}. */
int usg_mesh_cells(const usg_mesh_desc *mesh, int *cnode, int *cnv);
/* Free stream of usg_ctx_create for a Mach number and angle of attack
 * (TopologyHolderStrategy::set_fs_variable2, TopologyHolderStrategy.cpp:
 * 325-396): fs[4] = (rho, p, u, v). */
int usg_freestream_state(double mach, double aoa_deg, double *fs);
/* Tecplot FEBLOCK file of node coordinates and cell-centred density,
 * pressure, velocity, C_p = 2 (p - p_inf) / |u_inf|^2 and the reference's
 * Mach column |u| / sqrt(p / rho), named "<filename>_<mesh_tag+1>_<proc+1>"
 * -- TopologyHolderStrategy::writeCellData (TopologyHolderStrategyOutput.cpp:
 * 466-554), byte for byte. */
int usg_write_cell_data(const usg_mesh_desc *mesh, const double *U, const double *fs, const char *filename,
                        int proc, int mesh_tag);
/* Wall pressure coefficient -- Euler2D::outPutCp (Euler2D.cpp:392-623):
 * cell pressures averaged to the nodes, C_p = 2 (p - p_inf) / (rho_inf
 * |u_inf|^2), wall nodes with a signed count (negative where a wall face's
 * normal has a positive y component) sorted as rankWithXY (non-negative
 * counts by ascending x, then negative counts by descending x).
 * table[nwall*4] receives (x, y, count, C_p) of every wall node in that
 * order (the reference's HDF5 data set); curve[ncurve*4] the same after the
 * reference's merge of neighbours closer than 1e-4 in x (what its text file
 * plots).  table and curve hold `cap` rows; pass both NULL to get the counts
 * only. */
int usg_surface_cp(const usg_mesh_desc *mesh, const double *U, const double *fs, int cap, double *table,
                   double *curve, int *nwall, int *ncurve);
/* ---- overset connectivity (host only, no GPU needed) ---------------------
 * UniTioga's tioga::performConnectivity (UniTioga/tioga.cpp:152-212) for
 * nblocks 2-D blocks in one process, block i being mesh tag i+1: hole maps
 * from the wall faces (frc == USG_WALL), oriented boxes, the ADT donor
 * search, donor choice by cell resolution (area), fringes and cell iblanks.
 * overset_faces[i] (may be NULL, or NULL for a block) flags block i's faces
 * on its overset boundary (nf ints, nonzero = overset; the reader's
 * setOverSign faces, whose nodes become mandatory receptors).  mexclude
 * (cell layers next to the overset boundary that may not donate) and
 * nfringe (receptor layers around holes) are tioga::setMexclude /
 * setNfringe: MeshBlock's defaults are 3 and 1, the reference's input file
 * (test.input, read at Orchestra.cpp:105-110) sets 1 and 2.  Receptors
 * are nodes: their donor is a cell of another block and the stencil
 * weights its vertices (computeNodalWeight2D), as dataUpdate interpolates
 * node values (UniTioga/dataUpdate.cpp:6-133). */
typedef struct usg_overset usg_overset;
int usg_overset_create(const usg_mesh_desc *meshes, const int *const *overset_faces, int nblocks, int mexclude,
                       int nfringe, usg_overset **out);
int usg_overset_destroy(usg_overset *ov);
/* Node iblank[nn] (1 field, 0 hole, -1 receptor) and cell iblank[nc] (0 if a
 * vertex is a hole, -1 if every vertex is a receptor, else 1;
 * getCellIblanks, bookkeeping.cpp:654-705) of block `block`; either may be
 * NULL. */
int usg_overset_iblank(usg_overset *ov, int block, int *node_iblank, int *cell_iblank);
/* Receptors whose donor cell lies in block `block` (its INTERPLIST, in the
 * reference's order): *n receives their count; each array may be NULL.
 * recv_block/recv_node: the receptor; donor_cell: the donor cell of
 * `block` (input cell numbering); nw: 3 or 4 weights; donor_node[n*4] and
 * w[n*4]: the donor cell's vertices and weights (unused slots repeat the
 * first vertex with weight 0) -- the arrays usg_stencil_apply takes. */
int usg_overset_receptors(usg_overset *ov, int block, int *n, int *recv_block, int *recv_node, int *donor_cell,
                          int *nw, int *donor_node, double *w);
/* Resolutions after connectivity: node_res[nn], cell_res[nc] (cell area;
 * 1e15 marks cells excluded as donors and mandatory receptor nodes). */
int usg_overset_resolution(usg_overset *ov, int block, double *node_res, double *cell_res);
/* The connectivity's numerical building blocks, for tests: the partial
 * median partition of x[n] with companion ix[n] (find_median_wrap,
 * mathTools/median.cpp:188-232), and the oriented box of n points x[n*2]
 * (findOBB, UniTioga/utils.cpp:43-292): centre xc[2], half sizes dxc[2],
 * axes vec[2*2] (row i = axis i). */
int usg_overset_median(int *ix, double *x, int n);
int usg_overset_obb(const double *x, int n, double *xc, double *dxc, double *vec);

/* Cell-centred coupling of an overset assembly whose blocks are the cases
 * of a batch context (usg_batch_create, blocks in order): the donor stencil
 * of every receptor cell as one linear map of donor field cells, composed
 * of the stages the reference strings together around tioga::dataUpdate --
 * field-cell values averaged to the donor mesh's nodes (as Euler2D::outPutCp
 * averages cells to nodes, Euler2D.cpp:396-422), the nodal donor stencil
 * (MeshBlock::getInterpolatedSolution, UniTioga/dataUpdate.cpp:6-113) and
 * the receptor cell as the mean of its vertices (Orchestra.cpp:269-292).
 * Cells are numbered over the blocks in order (block b's cell c is
 * sum_{i<b} nc_i + c).  Outputs (host, caller allocated; pass recv NULL to
 * get the counts): blank[total cells] -- 1 field, 0 hole, -1 receptor (the
 * cell iblank of getCellIblanks; a receptor cell that has a vertex without
 * a donor, or donor nodes that touch no field cell, is left a field cell) --
 * nrecv receptors recv[nrecv], offsets off[nrecv+1], and ndonor donor cells
 * donor[ndonor] with weights w[ndonor] (ascending donor cell per receptor).
 * meshes are the blocks given to usg_overset_create.  Host only. */
int usg_overset_cell_stencil(usg_overset *ov, const usg_mesh_desc *meshes, int nblocks, int *blank, int *nrecv,
                             int *recv, int *off, int *ndonor, int *donor, double *w);

/* Overset coupling of a context (Euler or laminar flow; host arrays over
 * its cells, as usg_overset_cell_stencil returns them): every LU-SGS
 * iteration first sets each receptor cell's state from its donors,
 * U_r = sum_i w_i U_donor(i) and W_r = cons(U_r); the sweeps then update
 * field cells only (hole and receptor cells keep their state and take no
 * increment, so a field cell sees dW = 0 across the overset boundary), hole
 * and receptor cells reconstruct first order (a receptor's gradient reaches
 * into the hole cells beside it), and the residual norm sums field cells
 * only.  Receptors must be distinct
 * cells blanked -1 and donors cells blanked 1.  blank NULL removes the
 * coupling.  The reference registers node arrays with tioga and calls
 * dataUpdate from its driver (Orchestra.cpp:233-292); this is that step
 * inside the iteration. */
int usg_set_overset(usg_ctx *ctx, const int *blank, int nrecv, const int *recv, const int *off, const int *donor,
                    const double *w);

/* Pseudo-time CFL number of the LU-SGS iterations from the next iteration
 * on: the CFL argument of SolveTime (Euler2D.h:67-77), which the reference's
 * LUSGSIntegrator passes as 400 (LUSGS_Integrator.h:18).  A schedule that
 * starts low and grows lets stiff starts (wall spacing of y+ ~ 1) keep a
 * positive pressure. */
int usg_set_cfl(usg_ctx *ctx, double cfl);

/* Spatial order of the face reconstruction from the next residual on: 2 is
 * the reference's limited linear reconstruction (HLLCFluxStrategy.h:35-100,
 * the default), 1 uses the cell values (limiter phi = 0).  No counterpart
 * in the reference.  On meshes with y+ ~ 1 wall spacing the second-order
 * scheme diverged at every CFL number tried (Euler, laminar and SST); the
 * first-order one converges (SST on the 1M-cell airfoil: 4 decades). */
int usg_set_order(usg_ctx *ctx, int order);

/* Dual time stepping for unsteady flow (Euler and laminar; no counterpart
 * in the reference, whose iterations are pseudo-time only): with physical
 * step dt_phys > 0, each LU-SGS iteration solves the residual plus the
 * physical-time derivative of W by backward differences,
 *   R + V (c0 W + c1 W^n + c2 W^(n-1)),   D + c0 V,
 * with (c0, c1, c2) = (1, -1, 0) / dt_phys on the first physical step and
 * (3/2, -2, 1/2) / dt_phys after it when order is 2.  usg_dual_time_advance
 * starts a physical step (the current state becomes W^n, the previous W^n
 * becomes W^(n-1)); LU-SGS iterations are then its sub-iterations.  The
 * term is off until the first advance; dt_phys <= 0 switches it off. */
int usg_dual_time(usg_ctx *ctx, double dt_phys, int order);
int usg_dual_time_advance(usg_ctx *ctx);

/* The reference's "Cp_<step>" text file (Tecplot POINT zone of x and C_p,
 * the first point repeated at the end) of a curve from usg_surface_cp. */
int usg_write_cp(const char *path, const double *curve, int ncurve);

/* Plain device copy kernel (n = bytes/8 doubles) on a stream: the reference
 * kernel of known traffic used to calibrate HBM byte counters. */
int usg_copy_kernel(void *dst, const void *src, long bytes, void *hip_stream);

/* Last error text of the calling thread. */
const char *usg_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
