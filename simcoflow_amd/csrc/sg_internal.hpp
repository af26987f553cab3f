// Internal declarations of the simcogpu library (gfx950).
//
// Solver arrays use a padded layout with a one-cell zero ring: element (i,j)
// with i in [-1,ni], j in [-1,nj] lives at (j+2)*ld + i.  Rows start on a
// 256-byte boundary (ld is a multiple of 32 elements), the left ring cell of
// row j is the right padding of the storage row before it, and two extra
// storage rows hold the bottom ring.  Every kernel reads neighbours without
// bounds tests; the ring and padding stay zero because no kernel writes
// outside [0,ni)x[0,nj).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/simcogpu.h"

namespace sg {

constexpr double EPSI = 1.0e-3;  // StateVariables.f90:28
constexpr double TOLB = 1.0e-24; // ProjectionP_Mod.f90:404
constexpr int COARSEST_MAX = 16; // unknowns of the dense coarsest solve
constexpr int MAX_LEVELS = 24;
constexpr int NRED = 4;          // values a reduction slot can carry

// cells at least this much wider than tall coarsen in y first
constexpr double SEMI_ASPECT = 1.4;

struct Dims {
    int ni = 0, nj = 0, ld = 0;
    __host__ __device__ size_t alloc_elems() const { return (size_t)(nj + 3) * (size_t)ld; }
    __host__ __device__ size_t off(int i, int j) const {
        return (size_t)(j + 2) * (size_t)ld + (size_t)(i);
    }
    size_t interior_cells() const { return (size_t)ni * (size_t)nj; }
};

inline Dims make_dims(int ni, int nj) {
    Dims d;
    d.ni = ni;
    d.nj = nj;
    d.ld = ((ni + 1 + 31) / 32) * 32;
    return d;
}

// Operator of one level: symmetric 9-point stencil stored as positive
// couplings (matrix entry = -c).  Level 0 is 5-point: cNE/cNW are null.
template <typename T>
struct Op {
    T* diag = nullptr;
    T* cE = nullptr;  // (i,j)-(i+1,j)
    T* cN = nullptr;  // (i,j)-(i,j+1)
    T* cNE = nullptr; // (i,j)-(i+1,j+1)
    T* cNW = nullptr; // (i,j)-(i-1,j+1)
};

// Prolongation from level l+1 to level l, stored at the coarse resolution
// (coarse dims): for coarse point (I,J)
//   inj : fine (2I,2J)     <- (I,J)
//   xW,xE: fine (2I+1,2J)  <- (I,J),(I+1,J)
//   yS,yN: fine (2I,2J+1)  <- (I,J),(I,J+1)
//   SW,SE,NW,NE: fine (2I+1,2J+1) <- (I,J),(I+1,J),(I,J+1),(I+1,J+1)
template <typename T>
struct Prol {
    T* w[9] = {nullptr};
};
enum { W_INJ = 0, W_XW, W_XE, W_YS, W_YN, W_SW, W_SE, W_NW, W_NE };

template <typename T>
struct LevelVecs {
    T* x = nullptr;
    T* b = nullptr;
    T* r = nullptr;
};

struct Level {
    Dims d;
    Op<double> A;       // FP64 operator (setup is always FP64)
    Op<float> Af;       // FP32 copy (mixed precision)
    Prol<double> P;     // prolongation to this level from the next (if any)
    Prol<float> Pf;
    LevelVecs<double> v;
    LevelVecs<float> vf;
};

// Device-side scalars of the PCG iteration.
struct PcgScalars {
    double bb;      // ||b||^2 over fluid rows
    double rz;      // (r,z) of the current iteration
    double rz_old;  // (r,z) of the previous iteration
    double zq;      // (z_new, q_old) for the flexible beta
    double pq;      // (p,q)
    double rr;      // (r,r)
    double tol2;    // tol^2 * bb
    double alpha;   // step length of the current iteration
    double beta;    // direction update of the current iteration
    double true_rr; // ||b - A x||^2 after the solve
    double true_max;// max |b - A x| over fluid rows after the solve
    double mu;      // deflation: (A w, z) / (w, A w)
    double gamma;   // deflation: (w, A p) / (w, A w) of the last direction (re-projection)
    double mu_eff;  // deflation: coefficient of w in the next direction, mu + beta * gamma
    double wq;      // deflation: (w, A w)
    double wb;      // learned warm start: (w, b)
    double walpha;  // learned warm start: x0 = walpha * w
    double nn_scale;// rms of the restricted right-hand side (net input scale)
    double rr_ref;  // (r,r) at the last FP64 residual recomputation
    int active;     // 1 while not converged
    int iters;      // completed iterations
    int first;      // 1 before the first direction update
    int flexible;   // Polak-Ribiere beta (mixed precision)
    int guard;      // bit 0: NaN, bit 1: diag/rhs guard tripped
};

// work arrays of the device geometry rebuild (allocated at its first call)
struct GeoScratch {
    double* pfc[4] = {nullptr, nullptr, nullptr, nullptr};  // P-cell FCE (x, y), FCN (x, y), ghosted
    int* close[2] = {nullptr, nullptr};                     // closing-rule decisions, two passes
    int* changed = nullptr;
    int* moex[2] = {nullptr, nullptr};                      // U, V momentum-exchange cells
    double* next[2] = {nullptr, nullptr};                   // dx of column i+1, dy of row j+1
    double* old[4] = {nullptr, nullptr, nullptr, nullptr};  // aE, aN, vofSU, vofSV at the step's start
    bool have_old = false;
};

// ring of per-step diagnostics rows written on the device
constexpr int DIAG_BLOCKS = 1024;
struct DiagRing {
    double* rows = nullptr;       // [SG_DIAG_RING][SG_DIAG_COLS]
    long long* count = nullptr;   // rows written since the context was made
};
struct DiagState {
    DiagRing ring;
    double* partials = nullptr;   // [7][DIAG_BLOCKS]
    long long read = 0;           // rows handed to the host
};

struct Reducer {
    double* partials = nullptr;  // [NRED][max_blocks]
    unsigned int* tickets = nullptr;
    int max_blocks = 0;
};

// Work arrays of the interface transport (cell layout).
struct ClsBuffers {
    double* xc = nullptr;   // P-cell centres PGrid%x(:,1), PGrid%y(1,:)
    double* yc = nullptr;
    double* dis = nullptr;  // PLIC plane constants of the last reconstruction
    double* vt = nullptr;   // vof, phi after the first sweep of a sub-step
    double* pt = nullptr;
    double* tv = nullptr;   // temvf, temls of the first sweep (explicit term)
    double* tl = nullptr;
    unsigned char* stale = nullptr;    // plane constant left unassigned
    unsigned char* hit = nullptr;      // re-distancing: two-phase rule applied
    unsigned long long* best = nullptr;  // re-distancing: nearest distance (bits)
    int* srcs = nullptr;               // listed interface cells (re-distancing sources)
    int* stales = nullptr;             // listed cells with a stale plane constant
    int* ifc = nullptr;                // listed interface cells of a reconstruction and the cells
                                       // whose face fluxes need the PLIC geometry
    double* fr = nullptr;              // the latter's PLIC flux through the far / near face of a sweep
    double* fl = nullptr;
    unsigned* cnt = nullptr;           // list lengths, three per reconstruction
    static constexpr int CNT_SLOTS = 6 * 1024;
    unsigned long long* mask = nullptr;  // interface cells, 1 bit per cell
    // per mask word: any interface cell within the band of its 64 cells
    // (the OR over rows j-BAND..j+BAND of words w-1..w+1)
    unsigned long long* dil = nullptr;
    int* band = nullptr;               // the nonzero words of dil (w + j * mw)
    int mw = 0;                        // mask words per row
    double xw = 0.0;                   // west edge x of the grid (host copy)
    double dx0 = 0.0;                  // width of the first column (host copy)
};

// Staggered-cell geometry on the device: cell arrays first, then ghosted.
enum {
    STG_VOFS = 0, STG_NXS, STG_NYS, STG_PHIS, STG_WLLH, STG_DELH, STG_CCX, STG_CCY,
    STG_NCELL,
    STG_AE = STG_NCELL, STG_AW, STG_AN, STG_AS, STG_ETAE, STG_ETAN, STG_ALE, STG_ALN,
    STG_DALE, STG_DALN, STG_SXE, STG_SYN, STG_FCEX, STG_FCEY, STG_FCNX, STG_FCNY,
    STG_COUNT
};
struct StagDev {
    double* a[STG_COUNT] = {nullptr};
    int* moex = nullptr;
    int* act = nullptr;  // row of the momentum system
};

// Device scalars of the momentum BiCGStab.
struct BiScalars {
    double rho, rho_old, alpha, omega, beta, rv, ts, tt, rr, bb, tol2;
    int active;  // 1 while the Jacobi-scaled residual is above tol
    int iters;   // completed iterations
    int nan;     // the residual became NaN
};

// Work arrays of the momentum predictor.
struct PredBuffers {
    double *gu = nullptr, *gv = nullptr;   // FluxDiv before mixing
    double *fdu = nullptr, *fdv = nullptr; // FluxDiv
    double *mu = nullptr, *mv = nullptr;   // momentum rows (5 planes)
    double *ru = nullptr, *rv = nullptr;   // right-hand sides
    double *xu = nullptr, *xv = nullptr;   // solutions
    double* vec[8] = {nullptr};            // BiCGStab vectors
    double* partials = nullptr;            // 4 planes of BI_BLOCKS
    BiScalars* sc = nullptr;               // device
    BiScalars* sc_host = nullptr;          // pinned
    static constexpr int BI_BLOCKS = 2048;
};

namespace nn {
struct Net;
}

// Learned warm start / deflation vector of the projection PCG.
struct ParticleState {
    int n = 0, cap = 0;
    double *x = nullptr, *y = nullptr, *u = nullptr, *v = nullptr, *d = nullptr;
    unsigned* flagged = nullptr;
    sg_particle_params last{};  // parameters and dt of the last tracking call
    double last_dt = 0.0;
    bool tracked = false;
};

struct NnState {
    nn::Net* net = nullptr;  // inference net (batch 1) on level `level`
    int level = 0;
    int on = 0;
    int deflate = 0;
    double out_scale = 1.0;    // pressure = out_scale * rhs scale * net output
    void* X = nullptr;         // BF16 input [H+2][W+2][8]
    double* w = nullptr;       // fine deflation vector (padded FP64)
    double* aw = nullptr;      // A w
    float ms = 0.f;            // device time of the last warm start
    hipGraphExec_t graph = nullptr;  // the warm start's kernels, captured once
    hipGraphExec_t graph_prev = nullptr;  // the same on top of the previous solution
    // a caller's level-L prediction used in place of the net's output
    // (sg_set_learned_prediction): padded FP64 at the level's layout
    double* ext = nullptr;
    int ext_on = 0;
};

struct Ctx {
    int ni = 0, nj = 0;
    sg_params prm{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    // grid spacing
    double *dx = nullptr, *dy = nullptr, *udx = nullptr, *vdy = nullptr;
    // device fields in the reference layout
    void* field[SG_F_COUNT] = {nullptr};
    size_t field_bytes[SG_F_COUNT] = {0};
    // solver
    int nlev = 0;
    // level 1 coarsens level 0 in y only (cells at least 1.4x wider than
    // tall: the N couplings are 2x the E ones or more, which point
    // smoothing leaves to the coarse grid); all other levels coarsen 2x2
    int semi0 = 0;
    Level lev[MAX_LEVELS];
    double* coarse_inv = nullptr;  // dense inverse of the coarsest operator
    float* coarse_inv_f = nullptr;
    int coarse_n = 0;
    // PCG vectors on the fine level (FP64)
    double *x = nullptr, *r = nullptr, *z = nullptr, *p0 = nullptr, *p1 = nullptr, *q = nullptr;
    // FP32 PCG vectors of the mixed-precision solve (x and the RHS stay FP64)
    float *rf = nullptr, *zf = nullptr, *pf0 = nullptr, *pf1 = nullptr, *qf = nullptr;
    double* rhs = nullptr;
    double* stage = nullptr;  // cell-layout staging, 5*ni*nj
    int pcur = 0;
    PcgScalars* sc = nullptr;      // device
    PcgScalars* sc_host = nullptr; // pinned host mirror
    Reducer red;
    int built = 0;      // operator + rhs built
    double pcg_rate = 0.0;  // log (r,r) ratio per PCG iteration seen by the last solve (0: none yet)
    int mg_ready = 0;   // hierarchy set up for the current operator
    double vb = 0.0;
    // captured pair of PCG iterations, per precision (kernel arguments are
    // fixed for the life of the context, so one capture serves every solve)
    // [precision][0: two iterations from parity 0, 1: one from parity 0,
    // 2: one from parity 1]
    hipGraphExec_t iter_graph[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    int graph_pre[2] = {0, 0}, graph_post[2] = {0, 0}, graph_defl[2] = {0, 0}, graph_kry[2] = {0, 0};
    NnState nn;
    ParticleState part;
    double* step_dev = nullptr;   // device scalars of the time-step control (dt)
    DiagState diag;
    GeoScratch geo;
    // a solid moved inside the transport's sub-steps (kind NONE: static)
    struct {
        sg_solid_shape shape{};
        double us = 0.0, vs = 0.0, asx = 0.0, asy = 0.0;
    } moving;
    hipEvent_t ev_nn0 = nullptr, ev_nn1 = nullptr;
    // live timing of the transport's sweep launches (sg_sweep_timing): an
    // event pair per launch, drained into the totals when the pool fills
    struct SweepTiming {
        static constexpr int POOL = 512;  // event pairs
        int on = 0;
        int used = 0;
        hipEvent_t ev[2 * POOL] = {};
        long long launches = 0;
        double ms = 0.0;
    } sweep_timing;
    int use_graph = 1;
    ClsBuffers cls;
    StagDev stag[2];  // U, V cells
    PredBuffers pred;
    std::vector<void*> allocations;
};

// ---- launch helpers -------------------------------------------------------
constexpr int TX = 64;  // one wavefront along x
constexpr int TY = 4;
inline dim3 grid2d(int ni, int nj) { return dim3((ni + TX - 1) / TX, (nj + TY - 1) / TY); }
inline dim3 block2d() { return dim3(TX, TY); }

// Tile of this workgroup with the grid's workgroups dealt out to the chip's 8
// XCDs in contiguous bands: the hardware hands workgroup b to XCD b % 8, so
// tile start(b % 8) + b / 8 gives each XCD (and its L2) one band of
// neighbouring tiles, whose shared halo rows then hit in that L2.
__device__ inline void xcd_tile(int& tx, int& ty) {
    const int nb = (int)(gridDim.x * gridDim.y);
    const int b = (int)(blockIdx.x + blockIdx.y * gridDim.x);
    const int per = nb / 8, rem = nb % 8, x = b % 8;
    const int t = x * per + (x < rem ? x : rem) + b / 8;
    tx = t % (int)gridDim.x;
    ty = t / (int)gridDim.x;
}

// error handling
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);
#define SG_HIP(call)                                   \
    do {                                               \
        hipError_t _e = (call);                        \
        if (_e != hipSuccess) return hip_fail(_e, #call); \
    } while (0)

// ---- entry points of the kernel files (host wrappers) --------------------
// projection set-up, apply, scatter, correction (sg_poisson.hip)
int poisson_build(Ctx* c);
int poisson_rhs_only(Ctx* c);
int poisson_apply_fine(Ctx* c, const double* x, double* y);  // padded layout
int poisson_scatter(Ctx* c);
int poisson_matrix_rows(Ctx* c, double* d_matr);
int poisson_rhs_cells(Ctx* c, double* d_out);  // SetPoissonVectors' b, cell layout
int velocity_correct(Ctx* c);
int padded_to_cell(Ctx* c, const double* src, double* dst_cell);
int cell_to_padded(Ctx* c, const double* src_cell, double* dst);
// multigrid (sg_mg.hip)
int mg_allocate(Ctx* c);
int mg_setup(Ctx* c);
// z = M^{-1} r on the fine level with the PCG vectors of the context's
// precision (r, z, q in FP64; rf, zf, qf in mixed precision); the per-block
// partials of (r,z) and (z,q) land in `partials` (2 planes of max_blocks),
// *dots_done says whether they did (a one-level hierarchy has no fused dots)
int mg_vcycle_dots(Ctx* c, double* partials, int max_blocks, int* dots_done);
int mg_partial_blocks(Ctx* c);  // blocks that wrote partials in mg_vcycle_dots
int mg_launch_fine(Ctx* c, int which);  // measurement: 1 down, 2 up, 3 copy
// PCG (sg_pcg.hip)
int pcg_solve(Ctx* c, sg_solve_info* info);
int fine_residual_norms(Ctx* c, double* relres, double* maxrow);
// learned warm start (sg_learned.hip)
int learned_warm_start(Ctx* c, bool prev);   // x = Galerkin-scaled prolongated net output
int learned_features(Ctx* c, int level, void* X, float* t, float* mask, double* scale);
int learned_enable(Ctx* c, const float* w, int level, double out_scale, int deflate);
int learned_set_prediction(Ctx* c, int level, const double* e_host, int deflate);
int sweep_timing_drain(Ctx* c);  // sg_clsvof.hip
void learned_release(Ctx* c);
// time-step control and boundary layer (sg_step.hip)
int time_step_size(Ctx* c, double cfl, double vint, double uref, double vs, double asy, double* dt);
int closed_box_bc(Ctx* c);
int time_step_size_noalloc(int ni, int nj, const double* udx, const double* vdy, const double* u, const double* v,
                           double cfl, double Fr, double vint, double uref, double vs, double asy, double* dt_dev,
                           hipStream_t stream);
int geometry_rebuild(Ctx* c, const sg_solid_shape* shape);
int object_force(Ctx* c, double vs, double* force, double* clp, double* clf);
int geometry_snapshot(Ctx* c);
int solid_fraction_step(Ctx* c, const sg_solid_shape* shape);
int inter_new_var(Ctx* c, double vb);
int diagnostics_record(Ctx* c);
int diagnostics_read(Ctx* c, int max_rows, double* rows, int* n_rows, long long* n_lost);

// momentum predictor (sg_predictor.hip, sg_momentum.hip)
int predictor_allocate(Ctx* c);
int predictor_assemble(Ctx* c, double dt, double vb, int first_step);
int predictor_finish(Ctx* c);
int particles_track(Ctx* c, double dt, const sg_particle_params* p, int* flagged);
int particles_launch(Ctx* c, double dt, const sg_particle_params* p);  // no host sync
int particles_reserve(Ctx* c, int n);  // grows the particle buffers to n
int momentum_solve(Ctx* c, const double* m, const double* b, double* x, double tol, int max_iter, int* iters,
                   double* relres);
// interface transport (sg_clsvof.hip)
int clsvof_allocate(Ctx* c);
int clsvof_advect(Ctx* c, double dt, int nsub, const double* levels);
int interface_to_uv(Ctx* c);

void* dev_alloc(Ctx* c, size_t bytes);

}  // namespace sg
