/*
 * simcogpu -- MI355X (gfx950) projection step and interface transport for
 * simcoflow-style 2D cut-cell two-phase flow.  Plain C ABI: plain pointers
 * and sizes, no C++ or torch types, so a Fortran host binds it with
 * iso_c_binding and Python with ctypes.
 *
 * Array layout is the reference's Fortran layout (column-major, x fastest):
 *   cell array      (1:ni,1:nj)        ni*nj doubles, element (i,j) at (i-1)+(j-1)*ni
 *   ghosted array   (0:ni+1,0:nj+1)    (ni+2)*(nj+2) doubles, element (i,j) at i+j*(ni+2)
 * Host-pointer entry points copy in/out and return after the device work has
 * finished; the *_resident entry points work on the context's device fields.
 * sg_clsvof_advect_resident, sg_interface_to_uv_resident,
 * sg_velocity_correct_resident, sg_closed_box_bc and sg_field_copy enqueue
 * their work on the context's stream and return without waiting for it
 * (every call of a context uses that one stream, so they stay in order);
 * sg_synchronize, the downloads and the calls that return device results
 * (solves, predictor, time step) wait, and an error of the enqueued device
 * work surfaces there.  A caller that reads the fields through
 * sg_field_ptr on another stream calls sg_synchronize first.
 *
 * Return codes: 0 OK; negative = argument / HIP / allocation error;
 * positive = numerical (1 no convergence within max_iter, 2 NaN found,
 * 3 value guard tripped).  sg_last_error() gives the message.
 */
#ifndef SIMCOGPU_H
#define SIMCOGPU_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SG_OK 0
#define SG_ERR_ARG -1
#define SG_ERR_HIP -2
#define SG_ERR_ALLOC -3
#define SG_ERR_STATE -4
#define SG_NOT_CONVERGED 1
#define SG_NAN 2
#define SG_GUARD 3

#define SG_PREC_FP64 0  /* every solver kernel in FP64 */
#define SG_PREC_MIXED 1 /* FP32 V-cycle, FP64 outer residual / dots / coarsest solve */

typedef struct sg_ctx sg_ctx;

typedef struct sg_params {
    /* fluids: liquid/gas density and the reference density
     * (StateVariables.f90:26-27, Main.f90:148: Roref = row) */
    double rho_l, rho_g, rho_ref;
    /* kinematic viscosities, nu_ref, Reynolds and Froude numbers (predictor) */
    double nu_l, nu_g, nu_ref, Re, Fr;
    /* pressure boundary conditions W,E,S,N: 1 Neumann, 0 Dirichlet
     * (ProjectionP_Mod.f90:55,64 use 1,1,1,0) */
    int bc_w, bc_e, bc_s, bc_n;
    int precision;   /* SG_PREC_FP64 or SG_PREC_MIXED */
    int max_iter;    /* PCG iteration cap (ProjectionP_Mod.f90:108 uses 50) */
    double tol;      /* stop at ||b-Ap||_2 <= tol*||b||_2 */
    int mg_pre, mg_post; /* smoothing sweeps per level (ProjectionP_Mod.f90:126: 1); the
                            fused cycle kernels carry one each: only 1 is accepted */
    int warm_start;  /* 0: x0 = 0 (ProjectionP_Mod.f90:319); 1: x0 = previous solution;
                        2: learned warm start + deflation vector (sg_set_learned);
                        3: previous solution + the learned correction predicted from
                        its residual (Galerkin-scaled) */
    int check_every; /* PCG iterations between host convergence checks; 0 (default): set from
                        the observed convergence rate */
    int mg_coarse_krylov; /* flexible-CG iterations on the first level of at most 1024 points,
                             each preconditioned by the V-cycle below it (0: one V-cycle);
                             default 4 */
    int momentum_method;  /* predictor's momentum systems: 0 (default) Jacobi sweeps, with
                             Jacobi-BiCGStab if they have not converged in 40 sweeps;
                             1 Jacobi-BiCGStab */
} sg_params;

typedef struct sg_solve_info {
    int iterations;
    int status;
    double relres;           /* ||b-Ap||_2/||b||_2, recomputed in FP64 after the solve */
    double max_row_residual; /* max |(Ap-b)_i| over fluid rows (ProjectionP_Mod.f90:70-93) */
    double ms;               /* device time of the whole call */
    double setup_ms;         /* coefficient build + multigrid setup */
    double warm_ms;          /* learned warm start (net + prolongation + scaling), 0 without it */
    double warm_relres;      /* ||b-A x0||_2/||b||_2 of the starting guess */
    double warm_energy;      /* (w,b)^2/(w,Aw): the share x0'A x0 of the solution's energy x'A x = x'b
                                that the learned start captures (0 without it) */
} sg_solve_info;

/* fields held on the device by a context (sg_field_ptr / upload / download) */
enum sg_field {
    /* P-cell arrays (ni*nj) */
    SG_F_VOF = 0, SG_F_PHI, SG_F_VOFS, SG_F_NYS, SG_F_WLLH, SG_F_NXS, SG_F_PHIS,
    SG_F_NX, SG_F_NY, SG_F_MRES, SG_F_VOFSU, SG_F_VOFSV,
    /* ghosted arrays ((ni+2)*(nj+2)) */
    SG_F_U, SG_F_V, SG_F_P, SG_F_UPRED, SG_F_VPRED, SG_F_PP, SG_F_DPU, SG_F_DPV,
    SG_F_AE, SG_F_AW, SG_F_AN, SG_F_AS,
    /* int32 cell arrays (ni*nj): UCell/VCell Posnu/=-1 .and. MoExCell/=1 */
    SG_F_UFLAG, SG_F_VFLAG,
    /* momentum predictor, P-cell arrays (ni*nj): the P interface at time n
     * (PCellO%vof/phi/nx/ny), the U/V liquid fractions after the transport
     * (UCell/VCell%vof) and at time n (UCellO/VCellO%vof) */
    SG_F_VOFO, SG_F_PHIO, SG_F_NXO, SG_F_NYO, SG_F_UVOF, SG_F_VVOF, SG_F_UVOFO, SG_F_VVOFO,
    /* ghosted: Flux_n1(:,:,1) and Flux_n1(:,:,2) (Solver_Mod.f90:49-50) */
    SG_F_FLUXN1U, SG_F_FLUXN1V,
    /* U / V cell level set and normal (UCell/VCell%phi, nx, ny), ni*nj */
    SG_F_UPHI, SG_F_UNX, SG_F_UNY, SG_F_VPHI, SG_F_VNX, SG_F_VNY,
    SG_F_COUNT
};

const char* sg_last_error(void);
int sg_params_default(sg_params* p);

/* Context: grid spacings of the tensor-product grid of Mesh_Mod.f90:54-175,
 * dx[ni], dy[nj] of the P cells, udx[ni] of the U cells and vdy[nj] of the V
 * cells.  Allocates every device buffer once (no allocation afterwards). */
int sg_ctx_create(int ni, int nj, const double* dx, const double* dy,
                  const double* udx, const double* vdy, const sg_params* p,
                  sg_ctx** out);
int sg_ctx_destroy(sg_ctx* ctx);
int sg_set_params(sg_ctx* ctx, const sg_params* p);

/* device fields */
void* sg_field_ptr(sg_ctx* ctx, int field);
int sg_upload(sg_ctx* ctx, int field, const void* host);
int sg_download(sg_ctx* ctx, int field, void* host);
/* the same from / to a device buffer of the caller on the same GPU */
int sg_upload_device(sg_ctx* ctx, int field, const void* dev);
int sg_download_device(sg_ctx* ctx, int field, void* dev);
int sg_synchronize(sg_ctx* ctx);

/* Cut-cell geometry computed by Cutcell_Mod (Grid_Preprocess,
 * Cutcell_Mod.f90:18-32): P-cell vofS, nyS, WlLh, the ghosted P apertures
 * E/W/N/S, and the U/V flags and solid fractions used by the correction. */
int sg_set_geometry(sg_ctx* ctx, const double* vofS, const double* nyS,
                    const double* WlLh, const double* aE, const double* aW,
                    const double* aN, const double* aS, const int* uflag,
                    const int* vflag, const double* vofSU, const double* vofSV);

/* ---- projection (replaces ProjectionP_Mod.f90:26-99) --------------------- */
/* Builds the operator from the device fields VOF, PHI, DPU, DPV and the
 * geometry, and the RHS from UPRED, VPRED (ComputePossionMatrixCoefficient
 * :395-580, SetPoissonMatrix :137-273, SetPoissonVectors :275-354). */
int sg_poisson_build_resident(sg_ctx* ctx, double vb);
/* PCG + multigrid solve of the built system, then scatter/ghost fill into the
 * device field PP (DeltaPressureGetValues/BoundaryCondition :356-393,582-595). */
int sg_poisson_solve_resident(sg_ctx* ctx, sg_solve_info* info);
/* Drop-in body of PoissonEquationSolver: host in, host out.  u_star/v_star =
 * TPred%u/v, DpU/DpV = PU%Dp/PV%Dp (ghosted), vof/phi = PCell%vof/phi,
 * Pp_out = Proj%Pp (ghosted). */
int sg_poisson_solve(sg_ctx* ctx, const double* u_star, const double* v_star,
                     const double* DpU, const double* DpV, const double* vof,
                     const double* phi, double vb, double* Pp_out,
                     sg_solve_info* info);

/* test hooks: the assembled rows in the reference's matr layout
 * (ni,nj,5: S,W,diag,E,N; ProjectionP_Mod.f90:183-252), the RHS, and
 * y = A x with the device operator (x, y cell arrays, host pointers). */
int sg_poisson_matrix(sg_ctx* ctx, double* matr_out);
int sg_poisson_rhs(sg_ctx* ctx, double* rhs_out);
int sg_poisson_apply(sg_ctx* ctx, const double* x, double* y);

/* ---- velocity correction (replaces ComputePUV_Mod.f90:59-217,227-245) ---- */
/* device fields in: UPRED, VPRED, PP, DPU, DPV, VOF, PHI + geometry;
 * in/out: U, V, P; out: MRES. */
int sg_velocity_correct_resident(sg_ctx* ctx, double vb);

/* ---- momentum predictor (replaces Predictor_UV, PredictorUV_mod.f90:28-497) --
 * Geometry of the U (which = 0) or V (which = 1) cells, computed once by the
 * host's cut-cell set-up (NewCellFace, Cutcell_Mod.f90:399-725).  Cell arrays
 * (ni*nj): vofS, nxS, nyS, phiS, WlLh, delh, Cell_Cent x/y (ccx, ccy),
 * MoExCell (moex), active (1 where Posnu/=-1: a row of the momentum system;
 * the U cells of column ni and the V cells of row nj are never active).
 * Ghosted arrays ((ni+2)*(nj+2)): the four edge apertures, EtaE, EtaN, AlE,
 * AlN, DAlE, DAlN, SxE, SyN and FCE / FCN (x, y). */
typedef struct sg_stag_geometry {
    const double *vofS, *nxS, *nyS, *phiS, *WlLh, *delh, *ccx, *ccy;
    const int *moex, *active;
    const double *aE, *aW, *aN, *aS, *EtaE, *EtaN, *AlE, *AlN, *DAlE, *DAlN, *SxE, *SyN;
    const double *fcex, *fcey, *fcnx, *fcny;
} sg_stag_geometry;

typedef struct sg_predictor_info {
    int iterations_u, iterations_v;  /* iterations of the two systems: Jacobi sweeps, or BiCGStab
                                        iterations with momentum_method = 1 */
    double relres_u, relres_v;       /* final ||D^-1(b-Ax)||/||D^-1 b||, D = diag(A) */
    double ms;                       /* device time of the assembly and the two solves (the
                                        scatter that ends the call is enqueued, not waited for) */
} sg_predictor_info;

int sg_set_staggered_geometry(sg_ctx* ctx, int which, const sg_stag_geometry* g);
/* device fields in: U, V, VOFO, PHIO, NXO, NYO, UVOF, VVOF, UVOFO, VVOFO,
 * VOFS, NXS, NYS, PHIS + staggered geometry; in/out: FLUXN1U/V, DPU, DPV
 * (active cells and ghosts), UPRED, VPRED (rows 1..ni-1 / 1..nj-1 and the
 * open top row of v).  first_step: itt == 1 (first-order convective
 * fluxes, no Adams-Bashforth).  The two momentum systems are solved to
 * ||D^-1(b-Ax)|| <= 1e-13 ||D^-1 b|| (the reference runs 50 PCG iterations). */
int sg_predictor_resident(sg_ctx* ctx, double dt, double vb, int first_step, sg_predictor_info* info);
/* test hook: the assembled U and V rows (matr layout ni*nj*5: S,W,diag,E,N)
 * and right-hand sides of the last sg_predictor_resident call */
int sg_predictor_rows(sg_ctx* ctx, double* matr_u, double* rhs_u, double* matr_v, double* rhs_v);

/* ---- interface on the U / V cells (replaces Initial_ClsVofUV,
 * Clsvof_Mod.f90:402-710): device fields in VOF, PHI, NX, NY, VOFS, NXS, NYS,
 * PHIS and the staggered geometry (vofS, phiS of sg_set_staggered_geometry);
 * out UVOF, UPHI, UNX, UNY, VVOF, VPHI, VNX, VNY.  (The particle volume and
 * source sums of that routine belong to the particle model, not kept.) */
int sg_interface_to_uv_resident(sg_ctx* ctx);

/* ---- interface transport (replaces Coupled_LS_VOF, Clsvof_Mod.f90:730-959) --
 * nsub sub-steps of alternating x/y PLIC volume sweeps (X_Sweep/Y_Sweep
 * :987-1245) with level-set transport, reconstruction (Interface_Reconstruct
 * :1250-1321) and re-distancing (Redistance :1525-1708), for solids at rest
 * (the geometry of sg_set_geometry plus the solid line NXS/NYS/PHIS).
 * Needs ni, nj >= 4. */
typedef struct sg_clsvof_params {
    int nsub; /* sub-steps per call (nv = 6, Clsvof_Mod.f90:754) */
    /* west inflow of X_Sweep (:1057-1059): row j takes in liquid through its
     * west face while y_j - dy_j/2 < depth + amp*sin(kw*xw - omega*(time +
     * k*dt/nsub)) at sub-step k, xw the west edge of the grid
     * (Depthw, Amp0, kw, omew of StateVariables.f90:24-25) */
    double depth, amp, kw, omega;
} sg_clsvof_params;

int sg_clsvof_params_default(sg_clsvof_params* p);
/* P-cell centres PGrid%x(:,1), PGrid%y(1,:) (default: faces from 0 by the
 * spacings given to sg_ctx_create). */
int sg_set_centres(sg_ctx* ctx, const double* xc, const double* yc);
/* device fields in: U, V (TVar%u/v, masked by UFLAG/VFLAG as :778-797),
 * geometry; in/out: VOF, PHI; out: NX, NY (PCell%nx/ny). */
int sg_clsvof_advect_resident(sg_ctx* ctx, double dt, double time, const sg_clsvof_params* p);
/* Drop-in body of Coupled_LS_VOF: host in/out.  vof/phi = PCell%vof/phi
 * (in/out), nx/ny = PCell%nx/ny (out, may be NULL), u/v = TVar%u/v
 * (ghosted), time = NondiT. */
int sg_clsvof_advect(sg_ctx* ctx, double* vof, double* phi, double* nx, double* ny, const double* u,
                     const double* v, double dt, double time, const sg_clsvof_params* p);

/* ---- time-step control (replaces ComputeTimeStep, Solver_Mod.f90:321-344) -
 * dt = min over the cells of the reference's CFL bounds on the device fields
 * U, V (cfl; Fr of sg_params; vint, uref = TVar%Vint, TVar%Uref; vs, asy =
 * the solid's velocity and acceleration, 0 for solids at rest), a device
 * min-reduction; dt also stays on the device. */
int sg_time_step_size(sg_ctx* ctx, double cfl, double vint, double uref, double vs, double asy, double* dt);
/* The same CFL step without a context, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  No counterpart in the reference.  Every pointer
 * is DEVICE memory the caller owns: udx (ni), vdy (nj) = UGrid%dx(:,1),
 * VGrid%dy(1,:); u, v ghosted (ni+2)*(nj+2) as TVar%u, TVar%v; dt (one
 * double) receives the step.  Launches on `stream` (a hipStream_t, NULL =
 * the null stream) and returns without waiting.  Allocates nothing: the
 * per-block partial minima go to a 16 KiB array of the library in device
 * memory (one per GPU, held until the process ends), so calls that overlap
 * on one GPU race.  Rejects ni*nj outside [1, 2^31 - 1 - 256*2048]. */
int sg_time_step_size_noalloc(int ni, int nj, const double* udx, const double* vdy, const double* u,
                              const double* v, double cfl, double Fr, double vint, double uref, double vs,
                              double asy, double* dt, void* stream);
/* ghost layer of a closed box on U, V, P: the box form of
 * Boundary_Condition_Var (StateVariables.f90:118-211) -- pressure and
 * tangential velocity copied into the ghosts, P = 0 above the open top,
 * U = 0 on the west/east walls, V = 0 on the bottom wall */
int sg_closed_box_bc(sg_ctx* ctx);
/* device copy of field src onto field dst of the same shape (CopyNewCell) */
int sg_field_copy(sg_ctx* ctx, int dst, int src);

/* ---- learned warm start (no counterpart in the reference: it starts PCG
 * from x0 = 0, ProjectionP_Mod.f90:319) --------------------------------------
 * A 4-level convolutional encoder-decoder (16/32/64/128 channels, 3x3
 * kernels, additive skips, ReLU; BF16 weights and activations, FP32
 * accumulation on the matrix cores) runs on level L of the solve's
 * multigrid hierarchy.  Input channels at each level-L point: the
 * restricted right-hand side / its rms, the Galerkin couplings E, N, NE, NW
 * / the diagonal, the liquid fraction of the fine block, y and x in [0,1].
 * Output: the level-L pressure / (out_scale * rms).  PCG starts from the
 * prolongated, Galerkin-scaled prediction w and (deflate = 1, FP64 and
 * mixed precision) keeps its directions A-orthogonal to w; the stopping
 * test is unchanged. */
typedef struct sg_nn sg_nn;
size_t sg_nn_num_weights(void);
/* He-normal initial weights (host array of sg_nn_num_weights floats) */
int sg_nn_init_weights(float* w, unsigned long long seed);
/* a net for level grids of H x W (H % 8 == 0, W % 128 == 0), batches of up
 * to `batch` samples, on its own stream */
int sg_nn_create(int H, int W, int batch, sg_nn** out);
int sg_nn_destroy(sg_nn* n);
int sg_nn_set_weights(sg_nn* n, const float* w);  /* host FP32; resets Adam */
int sg_nn_get_weights(sg_nn* n, float* w);
/* x: device BF16 [nb][H+2][W+2][8] with a zero border; y: device FP32 [nb][H][W] */
int sg_nn_forward(sg_nn* n, const void* x, float* y, int nb);
/* loss = mean over the nb*H*W points of mask*(y - t)^2 (t, mask: device FP32
 * [nb][H][W]); sg_nn_gradient returns its gradient w.r.t. the weights
 * (host, sg_nn_num_weights floats), sg_nn_train_step applies one Adam step
 * (beta 0.9 / 0.999) with learning rate lr.  loss may be NULL. */
int sg_nn_gradient(sg_nn* n, const void* x, const float* t, const float* mask, int nb, float* grad, float* loss);
int sg_nn_train_step(sg_nn* n, const void* x, const float* t, const float* mask, int nb, float lr, float* loss);
/* grid of multigrid level `level` of a context (0 = the fine grid) */
int sg_nn_level_dims(sg_ctx* ctx, int level, int* H, int* W);
/* net input of the last solved system at `level` (device BF16 [H+2][W+2][8],
 * border left as is) and, when t != NULL, the training target t = solution
 * at the level's points / scale and the fluid mask (device FP32 [H][W]);
 * scale = rms of the restricted right-hand side. */
int sg_nn_features(sg_ctx* ctx, int level, void* x, float* t, float* mask, double* scale);
/* enable the learned warm start for solves with warm_start = 2: weights
 * (host FP32), the level it runs on, out_scale, deflate 0/1; w = NULL
 * disables it */
int sg_set_learned(sg_ctx* ctx, const float* w, int level, double out_scale, int deflate);
/* A level-L prediction from the caller in place of the net's output: the
 * pressure at the level's points (warm_start = 2), or the correction of the
 * previous pressure (warm_start = 3), which the solve prolongates, Galerkin-
 * scales and (deflate = 1) keeps its directions A-orthogonal to, exactly as
 * it does the net's.  No counterpart in the reference (its solves start from
 * zero); for callers with their own predictor, and to measure what any
 * predictor on that level could gain.  e: host array of nj_L x ni_L doubles,
 * row-major (sg_nn_level_dims gives the level's grid); NULL switches it
 * off.  Replaces a net set by sg_set_learned, and is replaced by one. */
int sg_set_learned_prediction(sg_ctx* ctx, int level, const double* e, int deflate);

/* ---- Lagrangian particles (replaces TrackingParticles, Particles.f90:71-205,
 * with Drag :206-223 and ParticlePosition :228-249) -------------------------
 * A particle set lives on the device with the context.  sg_particles_track
 * advances it over one time step dt (non-dimensional, as the flow) in
 * itp = 10 sub-steps of the reference's exponential drag integrator
 * (Schiller-Naumann drag, added mass Ca = 0.5, the fluid velocity
 * interpolated in the particle's P-cell at its position at the start of the
 * step), one-way coupled: it reads the device fields VOF, U, V, the geometry
 * (VOFS) and the P-cell centres, and changes nothing else.  flagged: the
 * number of particles that met the reference's failure test (NaN velocity,
 * |x| > 1e10 or |ax| > 1e10; the reference pauses). */
typedef struct sg_particle_params {
    double rho_l, rho_g; /* row, roa (StateVariables.f90:26-27) */
    double nu_l, nu_g;   /* nuw, nua */
    double rho_p;        /* rop, the particles' density (Main.f90:150) */
    double gx, gy;       /* gravity components as the reference's gx, gy (Main.f90:200-201) */
    double lref, uref;   /* PGrid%Lref, TVar%Uref: positions in metres, velocities in m/s */
    double us, vs;       /* velocity of the solid (BoomCase%us, %vs) */
} sg_particle_params;

/* positions x, y, velocities u, v, diameters d of n particles (host) */
int sg_particles_set(sg_ctx* ctx, int n, const double* x, const double* y, const double* u, const double* v,
                     const double* d);
int sg_particles_get(sg_ctx* ctx, int n, double* x, double* y, double* u, double* v);
int sg_particles_track(sg_ctx* ctx, double dt, const sg_particle_params* p, int* flagged);

/* ---- cut-cell geometry rebuilt on the device (replaces Grid_Preprocess,
 * Cutcell_Mod.f90:18-725, with SolidVolumeFraction, Clsvof_Mod.f90:2553-2929,
 * for a solid given as a shape; ObjectMovement, Clsvof_Mod.f90:2384-2398, is
 * the host's: it moves the shape's centre before each call) -----------------
 * The solid level set and normal on the P, U and V cells (a signed distance,
 * positive in the fluid), the solid fractions, apertures (harmonic face
 * averaging, solid and closed cells shut), wall lengths, the U/V solid
 * fractions and flags, the momentum-exchange cells and every staggered-cell
 * array of sg_set_staggered_geometry: the device state sg_set_geometry +
 * sg_set_staggered_geometry would set, computed on the device.  box: half
 * width a, half height b; circle: radius a; ellipse: semi-axes a, b; boom
 * (the oil boom BoomCase of Main.f90:173-182, moved by ObjectMovement,
 * Clsvof_Mod.f90:2388-2398): a cylinder of radius a (Dobj/2) with a bar of
 * half width b (Wobj/2 < a) reaching c (LBar) below where it leaves the
 * cylinder, as the distance to the union of the two; none: no solid.  Needs ni, nj >= 4.  The first call allocates the rebuild's work
 * arrays (4 ghosted + 2 cell planes); later calls allocate nothing and wait
 * on the host once per pass of the closing rule (1-3 passes). */
#define SG_SOLID_NONE 0
#define SG_SOLID_BOX 1
#define SG_SOLID_CIRCLE 2
#define SG_SOLID_ELLIPSE 3
#define SG_SOLID_BOOM 4
typedef struct sg_solid_shape {
    int kind;
    double cx, cy; /* centre */
    double a, b, c;
} sg_solid_shape;
int sg_geometry_rebuild(sg_ctx* ctx, const sg_solid_shape* shape);
/* force on the solid (replaces ComputeForceObject, Clsvof_Mod.f90:2400-2549,
 * with ICorProb = .TRUE., Main.f90:205): the vertical force clf - clp from
 * the device fields P, V, the P geometry (VOFS NYS WLLH AE AW) and the V-cell
 * geometry (vofS, delh, moex), with Rey = sg_params.Re and the solid's
 * vertical velocity vs; clp (pressure part) and clf (shear part) may be
 * NULL.  The host integrates ObjectMovement with it (asy = (force - M g)/M,
 * Solver_Mod.f90:277-283). */
int sg_object_force(sg_ctx* ctx, double vs, double* force, double* clp, double* clf);
/* a solid that moves inside the transport (ObjectMovement + SolidVolumeFraction
 * on the P cells after each sub-step's sweeps, Clsvof_Mod.f90:920-921): once
 * armed with a shape (kind != NONE), velocity (us, vs) and acceleration
 * (asx, asy), every sub-step of sg_clsvof_advect_resident / sg_clsvof_advect
 * moves the shape by ObjectMovement's kinematics (:2384-2398) and recomputes
 * the P cells' PHIS, NXS, NYS, VOFS for it (snapped below epsi and above
 * 1 - epsi) and carries VOF onto the new fluid part of each cell
 * (:2906-2917).  sg_get_moving_solid reads the moved shape and velocity
 * back; kind NONE disarms.  The U/V cells and apertures follow with
 * sg_geometry_rebuild after the transport, as Grid_Preprocess does.  Both
 * calls reject (SG_ERR_ARG) an unknown kind, a centre that is not finite,
 * sizes that are not positive and a boom whose bar is not narrower than its
 * float (b >= a) or has c < 0; sg_set_moving_solid also a velocity or
 * acceleration that is not finite.  The shape (and motion) checks come before
 * the context is used. */
int sg_set_moving_solid(sg_ctx* ctx, const sg_solid_shape* shape, double us, double vs, double asx, double asy);
int sg_get_moving_solid(sg_ctx* ctx, sg_solid_shape* shape, double* us, double* vs);
/* the velocities after the solid moved (replaces InterNewVar,
 * Solver_Mod.f90:360-495): sg_geometry_snapshot keeps the geometry of the
 * step's start (the part CopyNewCell's PCellO/UCellO/VCellO give it: the P
 * apertures AE, AN and the U/V solid fractions; its first call allocates
 * them); after sg_geometry_rebuild, sg_inter_new_var sets u = 0 on U faces
 * whose cell the solid uncovered, v = vb on such V faces, and scales u / v
 * on P faces whose open fraction shrank below 0.6 of its old value (either
 * above 0.5) by old/new.  The reference's other computations there are
 * unused and left out. */
int sg_geometry_snapshot(sg_ctx* ctx);
int sg_inter_new_var(sg_ctx* ctx, double vb);
/* test hook: one staggered-geometry array of U (which = 0) or V cells, in the
 * order of sg_stag_geometry's double arrays (vofS nxS nyS phiS WlLh delh ccx
 * ccy: ni*nj; aE aW aN aS EtaE EtaN AlE AlN DAlE DAlN SxE SyN fcex fcey fcnx
 * fcny: (ni+2)*(nj+2)), then moex and active (int, ni*nj) */
int sg_stag_download(sg_ctx* ctx, int which, int k, void* host);

/* ---- per-step diagnostics without host synchronisation -------------------
 * sg_diagnostics_record enqueues, on the context's stream, two kernels that
 * reduce the current device fields (deterministic two-phase sums and
 * maxima) into one row of a device ring of SG_DIAG_RING rows; nothing waits
 * for them.  Row columns:
 *   0 dt of the last sg_time_step_size (ComputeTimeStep, Solver_Mod.f90:321)
 *   1 liquid volume  sum vof dx dy
 *   2 gas volume     sum (1 - vof - vofS) dx dy
 *   3, 4 gas centroid (x, y), gas-volume weighted cell centres
 *   5 max |mres|: the mass residual of the last velocity correction
 *     (ComputePUV_Mod.f90:184-217)
 *   6, 7 max |u|, max |v| over the faces U(1:ni,1:nj), V(1:ni,1:nj)
 * (maxima propagate NaN).  sg_diagnostics_read waits for the stream and
 * copies the oldest rows not yet read, at most max_rows, in recording
 * order; n_lost (may be NULL) counts rows overwritten before they were
 * read.  The reference writes its per-step history from the host
 * (PrintHistory, Solver_Mod.f90:346-356). */
#define SG_DIAG_COLS 8
#define SG_DIAG_RING 4096
int sg_diagnostics_record(sg_ctx* ctx);
int sg_diagnostics_read(sg_ctx* ctx, int max_rows, double* rows, int* n_rows, long long* n_lost);

/* ---- checkpoint (replaces the Tecplot-ASCII restart of RunAgain,
 * Solver_Mod.f90:164-195 with ReadOldData*, PrintResult.f90:192-270) -------
 * The state is what a time step reads from the one before it: the P and U/V
 * interfaces (VOF PHI NX NY, UVOF VVOF UPHI UNX UNY VPHI VNX VNY), U V P,
 * FLUXN1U FLUXN1V, DPU DPV (an inactive U/V cell keeps its last Dp, which
 * the next Poisson assembly can read), the Poisson solution and the two
 * momentum solutions (the starts of the next solves) and the particle set,
 * in FP64 as held, so a run
 * resumed from it repeats the uninterrupted run bit for bit.  Not included:
 * the geometry (set it up first, as the reference recomputes it), the
 * parameters, and the host's step counter and time.  sg_state_bytes gives
 * the size of the current state; sg_state_load takes a state saved from a
 * context of the same grid (grows the particle buffers if needed). */
int sg_state_bytes(sg_ctx* ctx, size_t* bytes);
int sg_state_save(sg_ctx* ctx, void* buf, size_t bytes);
int sg_state_load(sg_ctx* ctx, const void* buf, size_t bytes);

/* ---- measurement ----------------------------------------------------------
 * Average device time (HIP events on the context's stream) of `reps`
 * back-to-back launches of one kernel on the current data: which = 0 the
 * fine-level matrix-free apply y = A x, 1 the fine-level multigrid
 * down-sweep, 2 the fine-level up-sweep (with the fused PCG dots), 3 a plain
 * copy of one padded fine-level vector (counter calibration), 4 the particle
 * tracker on the current particle set with the last sg_particles_track call's
 * dt and parameters (the set advances reps + 1 steps), 5 sg_diagnostics_record
 * (writes reps + 1 rows).  0 needs the operator built, 1-3 one solve, 4 one
 * tracking call.  No counterpart in the reference (it has no timers). */
int sg_time_kernel(sg_ctx* ctx, int which, int reps, double* avg_ms);
/* Live timing of the interface transport's sweep launches (X_Sweep /
 * Y_Sweep of Coupled_LS_VOF): on = 1 clears the totals and brackets every
 * sweep launch of the calls that follow with HIP events on the context's
 * stream; on = 0 waits for the recorded launches and returns their number and
 * summed duration.  No counterpart in the reference (measurement). */
int sg_sweep_timing(sg_ctx* ctx, int on, long long* launches, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif
