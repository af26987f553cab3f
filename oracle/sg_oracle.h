/*
 * CPU FP64 oracle for the simcoflow projection step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path may link or call
 * this library: it is the checker used by tests/, the smoke check and the
 * CPU-baseline leg of bench.py.
 *
 * Layout (identical to the reference's Fortran arrays, column-major, x fastest):
 *   cell arrays     (1:ni,1:nj)        a[(i-1) + (j-1)*ni]
 *   ghosted arrays  (0:ni+1,0:nj+1)    a[i + j*(ni+2)]
 *   PoCoef          (1:ni,1:nj,1:4)    faces S,W,E,N   (ProjectionP_Mod.f90:45)
 *   matr            (1:ni,1:nj,1:5)    S,W,diag,E,N    (ProjectionP_Mod.f90:183-252)
 *
 * Each function names the reference lines it restates.
 */
#ifndef SG_ORACLE_H
#define SG_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Wet/dry classification used by ProjectionP_Mod.f90:409-410,493-494 and
 * ComputePUV_Mod.f90:63-64,80-81.  Returns 1 wet, 0 dry. */
int orc_is_wet(double vof, double phi, double vofS);

/* ComputePossionMatrixCoefficient, ProjectionP_Mod.f90:395-580. */
void orc_poisson_coef(int ni, int nj, const double* dx, const double* dy,
                      const double* udx, const double* vdy,
                      const double* vof, const double* phi, const double* vofS,
                      const double* DpU, const double* DpV,
                      double rho_l, double rho_g, double rho_ref,
                      double* PoCoef);

/* SetPoissonMatrix, ProjectionP_Mod.f90:137-273 (row values only; bc flags
 * 1 = Neumann, 0 = Dirichlet).  Solid rows (vofS >= 1-epsi) are left zero. */
void orc_poisson_matrix(int ni, int nj, const double* PoCoef,
                        const double* aE, const double* aW,
                        const double* aN, const double* aS,
                        const double* vofS, int bcW, int bcE, int bcS, int bcN,
                        double* matr);

/* SetPoissonVectors, ProjectionP_Mod.f90:275-354.  rhs on fluid rows, 0 on
 * solid rows. */
void orc_poisson_rhs(int ni, int nj, const double* dx, const double* dy,
                     const double* u, const double* v,
                     const double* aE, const double* aW,
                     const double* aN, const double* aS,
                     const double* nyS, const double* WlLh, const double* vofS,
                     double vb, double* rhs);

/* Row-wise product y = A x exactly as the residual check of
 * ProjectionP_Mod.f90:70-85 forms it (x is a cell array). */
void orc_matr_apply(int ni, int nj, const double* matr, const double* x,
                    double* y);

/* max over fluid rows of |(A x)_i - b_i|, ProjectionP_Mod.f90:70-93. */
double orc_max_row_residual(int ni, int nj, const double* matr,
                            const double* x, const double* b,
                            const double* vofS);

/* DeltaPressureGetValues + DeltaPressureBoundaryCondition,
 * ProjectionP_Mod.f90:356-393,582-595: scatter the cell solution into the
 * ghosted Pp (solid cells -> 0) and fill ghosts. */
void orc_pressure_scatter_bc(int ni, int nj, const double* x,
                             const double* vofS, int bcW, int bcE, int bcS,
                             int bcN, double* Pp);

/* Velocity correction + top-row continuity + mass residual + internal-cell
 * condition, ComputePUV_Mod.f90:59-217,227-245.  u,v,p are the ghosted
 * TVar%u/v/p (in/out), upred/vpred the ghosted predictor velocities,
 * Pp the ghosted projection pressure.  uflag/vflag: 1 where
 * UCell%Posnu/=-1 and UCell%MoExCell/=1 (resp. V), else 0. */
void orc_velocity_correct(int ni, int nj, const double* dx, const double* dy,
                          const double* udx, const double* vdy,
                          const double* vof, const double* phi,
                          const double* vofS, const double* nyS,
                          const double* WlLh,
                          const double* aE, const double* aW,
                          const double* aN, const double* aS,
                          const int* uflag, const int* vflag,
                          const double* vofSU, const double* vofSV,
                          const double* DpU, const double* DpV,
                          const double* upred, const double* vpred,
                          const double* Pp, double vb,
                          double rho_l, double rho_g, double rho_ref,
                          double* u, double* v, double* p, double* mres);

/* ---- CPU restatement of the solver (the CPU baseline) -----------------
 * PCG preconditioned by the same structured geometric multigrid V-cycle the
 * GPU path runs (2x2 aggregation, Galerkin coarse operators, red-black
 * Gauss-Seidel, direct coarsest solve).  The operator is the symmetric form
 * of matr: cE(i,j) = -matr(i,j,E), cN(i,j) = -matr(i,j,N), diag = matr(.,3)
 * (0 on solid rows).  Stops when ||b-Ax||_2 <= tol*||b||_2.
 * Returns the iteration count (negative if max_iter was reached). */
int orc_pcg_mg_solve(int ni, int nj, const double* matr, const double* b,
                     double* x, double tol, int max_iter, int pre, int post,
                     double* relres_out);

/* Plain Jacobi-PCG on the same operator (reference iteration for tests). */
int orc_pcg_jacobi_solve(int ni, int nj, const double* matr, const double* b,
                         double* x, double tol, int max_iter,
                         double* relres_out);

/* Number of OpenMP threads the solver uses (1 when built without OpenMP). */
int orc_num_threads(void);
void orc_set_num_threads(int n);

/* ---- coupled level-set / VOF transport (Clsvof_Mod.f90) ------------------ */

/* Grid and static solid geometry seen by Coupled_LS_VOF: dx[ni], dy[nj]
 * spacings and xc[ni], yc[nj] centres of the P cells (PGrid%dx/x), the
 * P-cell solid fraction and solid line (vofS, nxS, nyS, phiS; cell arrays),
 * the ghosted apertures E/W/N/S and the U/V face flags (cell arrays, 1 where
 * Posnu/=-1 .and. MoExCell/=1). */
typedef struct orc_cls_geom {
    int ni, nj;
    const double *dx, *dy, *xc, *yc;
    const double *vofS, *nxS, *nyS, *phiS;
    const double *aE, *aW, *aN, *aS;
    const int *uflag, *vflag;
} orc_cls_geom;

/* West inflow wave of X_Sweep (:1057-1059): liquid enters row j while
 * yc_j - dy_j/2 < depth + amp*sin(kw*xw - omega*t), xw the west edge. */
typedef struct orc_wave {
    double depth, amp, kw, omega;
} orc_wave;

double orc_frac(double nx, double ny, double diss, double dx, double dy);
double orc_plic_distance(double vf, double nxx, double nyy, double dx, double dy);
void orc_cellgeocal(double nxs, double nys, double phis, double nxl, double nyl, double phil,
                    double dx, double dy, double* vols, double* volf);
int orc_distance_fluid(double nxss, double nyss, double phiss, double nxf, double nyf, double volf,
                       double dx, double dy, double* phif);
void orc_normal_irre(const orc_cls_geom* g, const double* phi, const double* vofS, int i, int j,
                     double* nx, double* ny);
void orc_reconstruct(const orc_cls_geom* g, const double* vof, const double* phi, double* nx, double* ny,
                     double* dis);
void orc_masked_velocities(const orc_cls_geom* g, const double* u, const double* v, double* ue, double* ve);
void orc_x_sweep(const orc_cls_geom* g, const double* vfl, const double* phi, const double* nx, const double* ny,
                 const double* dis, const double* ue, double dtv, double inflow_level, double* temvf,
                 double* temls);
void orc_y_sweep(const orc_cls_geom* g, const double* vfl, const double* phi, const double* nx, const double* ny,
                 const double* dis, const double* ve, double dtv, double* temvf, double* temls);
void orc_boundary_copy(int ni, int nj, double* a);
void orc_redistance(const orc_cls_geom* g, const double* vfl, const double* nxx1, const double* nyy1,
                    const double* diss1, double* phi);
double orc_inflow_level(const orc_cls_geom* g, const orc_wave* w, double time, int k, double dtv);
/* Initial_ClsVofUV (Clsvof_Mod.f90:402-710) for U (which=0) or V (which=1)
 * cells from the P-cell vof, phi, nx, ny; vofST / phiST the staggered
 * cells' solid fraction and solid level set; outputs are cell arrays. */
void orc_interface_to_uv(const orc_cls_geom* g, int which, const double* vof, const double* phi, const double* nx,
                         const double* ny, const double* vofST, const double* phiST, double* tvof, double* tphi,
                         double* tnx, double* tny);
/* nv sub-steps of Coupled_LS_VOF over dt; u, v ghosted TVar%u/v;
 * vof, phi in/out; nx, ny out. */
void orc_clsvof_advect(const orc_cls_geom* g, double* vof, double* phi, double* nx, double* ny,
                       const double* u, const double* v, double dt, double time, int nv, const orc_wave* w);

/* ---- momentum predictor (PredictorUV_mod.f90) ------------------------------ */

/* One staggered grid (U or V cells).  Cell arrays (1:ni,1:nj): vofS, nxS,
 * nyS, phiS, WlLh, delh, Cell_Cent (ccx, ccy), MoExCell (moex), the active
 * flag (Posnu/=-1: a row of the momentum system), the vof at time n (the
 * U/VCellO%vof) and after the transport (U/VCell%vof).  Ghosted arrays
 * (0:ni+1,0:nj+1): edge apertures, EtaE/N, AlE/N, DAlE/N, SxE, SyN and the
 * face centres FCE (fcex, fcey) and FCN (fcnx, fcny). */
typedef struct orc_stag {
    const double *vofS, *nxS, *nyS, *phiS, *WlLh, *delh, *ccx, *ccy;
    const int *moex, *flag;
    const double *vof_old, *vof_new;
    const double *aE, *aW, *aN, *aS;
    const double *EtaE, *EtaN, *AlE, *AlN, *DAlE, *DAlN, *SxE, *SyN;
    const double *fcex, *fcey, *fcnx, *fcny;
} orc_stag;

typedef struct orc_pred_in {
    int ni, nj;
    const double *dx, *dy, *udx, *vdy;  /* P spacings, U-cell dx, V-cell dy */
    orc_stag U, V;
    /* P cells at time n (PCellO) */
    const double *pvof, *pphi, *pnx, *pny, *pvofS, *pnxS, *pnyS, *pphiS;
    const double *u, *v;                /* ghosted TVar%u, TVar%v */
    double rho_l, rho_g, rho_ref, nu_l, nu_g, nu_ref, Re, Fr;
    double dt, vb;
    int first_step;                     /* itt == 1 */
} orc_pred_in;

/* Predictor_UV up to the two linear systems (PredictorUV_mod.f90:28-497):
 * U and V momentum rows (matr layout, 5 planes S,W,diag,E,N of ni*nj),
 * right-hand sides (ni*nj), PU%Dp / PV%Dp on active cells (ghosted, in/out)
 * and Flux_n1 (ghosted, 2 planes, in/out). */
void orc_predictor_assemble(const orc_pred_in* in, double* fluxn1, double* matr_u, double* rhs_u,
                            double* matr_v, double* rhs_v, double* DpU, double* DpV);
/* test hooks: the superbee limiter (:2354-2367), the high-order convective
 * face fluxes (:1186-1429; CF1 = CFluxEW, CF2 = CFluxNS, 2 ghosted planes
 * each) and Mixing_Procedure (:2127-2197) of U (comp 0) or V cells */
double orc_muscl_limiter(double r, double tolim);
void orc_conv_fluxes(const orc_pred_in* in, double* CF1, double* CF2);
void orc_mixing(const orc_pred_in* in, int comp, double* flux);
/* ghost copies of PU%Dp / PV%Dp (:530-540) and Pred%v(:,nj) = Pred%v(:,nj-1)
 * (PredictorVelocityBoundaryCondition :2199-2213) */
void orc_predictor_finish(int ni, int nj, double* DpU, double* DpV, double* vpred);

/* ---- time-step control (Solver_Mod.f90) ---------------------------------- */
double orc_time_step(int ni, int nj, const double* udx, const double* vdy, const double* u, const double* v,
                     double cfl, double Fr, double vint, double uref, double vs, double asy);
void orc_closed_box_bc(int ni, int nj, double* u, double* v, double* p);

/* sg_oracle_force.c -- ComputeForceObject (Clsvof_Mod.f90:2400-2549) */
double orc_object_force(int ni, int nj, const double* dx, const double* vofS, const double* nyS, const double* WlLh,
                        const double* aE, const double* aW, const double* p, const double* vofSV, const int* moexV,
                        const double* delhV, const double* v, double vs, double Rey, double* clp_out,
                        double* clf_out);

void orc_inter_new_var(int ni, int nj, const double* vofSUo, const double* vofSU, const double* vofSVo,
                       const double* vofSV, const double* aEo, const double* aE, const double* aNo,
                       const double* aN, double vb, double* u, double* v);

/* sg_oracle_particles.c -- Lagrangian particles (Particles.f90:71-249) */
typedef struct {
    double rho_l, rho_g;  /* row, roa */
    double nu_l, nu_g;    /* nuw, nua */
    double rho_p;         /* rop */
    double gx, gy;        /* gravity components */
    double lref, uref;    /* PGrid%Lref, TVar%Uref */
    double us, vs;        /* BoomCase%us, %vs */
} orc_particle_params;
double orc_particle_drag(double re);
void orc_particle_position(int ni, int nj, const double* xc, const double* yc, const double* dx, const double* dy,
                           double lref, double x, double y, int* ii, int* jj);
int orc_particle_track(int ni, int nj, const double* xc, const double* yc, const double* dx, const double* dy,
                       const double* vof, const double* vofS, const double* u, const double* v, int n, double* px,
                       double* py, double* pu, double* pv, const double* pd, const orc_particle_params* P,
                       double dt);

/* sg_oracle_cutcell.c -- cut-cell geometry (Cutcell_Mod.f90) */
typedef struct orc_cgrid {
    int ni, nj;
    const double *x, *y;   /* TGrid%x(i,.) per i, TGrid%y(.,j) per j */
    const double *dx, *dy; /* TGrid%dx per i, TGrid%dy per j */
} orc_cgrid;
/* One grid's Cell fields (Mesh_Mod.f90:11-41).  Cell arrays: vofS, phiS,
 * nxS, nyS, WlLh, delh, Cell_Cent (ccx, ccy), MoExCell (moex), Posnu.
 * Ghosted: apertures, FCE (fcex, fcey), FCN (fcnx, fcny) and the face
 * factors.  P cells: moex, delh, ccx/ccy and the face factors NULL. */
typedef struct orc_ccell {
    double *vofS, *phiS, *nxS, *nyS, *WlLh, *delh, *ccx, *ccy;
    int *moex, *posnu;
    double *aE, *aW, *aN, *aS, *fcex, *fcey, *fcnx, *fcny;
    double *EtaE, *EtaN, *AlE, *AlN, *DAlE, *DAlN, *SxE, *SyN;
} orc_ccell;
void orc_edge_geo_cal(double p1x, double p1y, double p2x, double p2y, double nxx, double nyy, double diss,
                      double* area, double* fx, double* fy);
void orc_solid_fraction(int ni, int nj, const double* dx, const double* dy, const double* phiS, const double* nxS,
                        const double* nyS, double* vofS);
void orc_cell_geo_cal(const orc_cgrid* G, orc_ccell* C, int icorprob);
void orc_define_moex(int ni, int nj, orc_ccell* P, orc_ccell* U, orc_ccell* V);
void orc_number_external(int ni, int nj, orc_ccell* C, int iu, int iv);
void orc_new_cell_face(const orc_cgrid* PG, const orc_cgrid* UG, const orc_cgrid* VG, orc_ccell* P, orc_ccell* U,
                       orc_ccell* V);
void orc_cut_cell_geometry(const orc_cgrid* PG, const orc_cgrid* UG, const orc_cgrid* VG, orc_ccell* P,
                           orc_ccell* U, orc_ccell* V, int icorprob);

#ifdef __cplusplus
}
#endif
#endif
