// C ABI of the simcogpu library: context, device fields, projection entry
// points.  Every device buffer is allocated in sg_ctx_create; the solve and
// correction entry points allocate nothing.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "sg_internal.hpp"
#include "sg_nn.hpp"

namespace sg {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what) {
    g_err = std::string("HIP error '") + hipGetErrorString(e) + "' in " + what;
    return SG_ERR_HIP;
}

void* dev_alloc(Ctx* c, size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) return nullptr;
    if (hipMemset(p, 0, bytes ? bytes : 16) != hipSuccess) {
        (void)hipFree(p);
        return nullptr;
    }
    c->allocations.push_back(p);
    return p;
}

static bool is_cell_field(int f) {
    return (f >= SG_F_VOF && f <= SG_F_VOFSV) || (f >= SG_F_VOFO && f <= SG_F_VVOFO) ||
           (f >= SG_F_UPHI && f <= SG_F_VNY);
}
static bool is_ghost_field(int f) { return (f >= SG_F_U && f <= SG_F_AS) || f == SG_F_FLUXN1U || f == SG_F_FLUXN1V; }
static bool is_int_field(int f) { return f == SG_F_UFLAG || f == SG_F_VFLAG; }

static void destroy(Ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& per : c->iter_graph)
        for (hipGraphExec_t g : per)
            if (g) (void)hipGraphExecDestroy(g);
    learned_release(c);
    if (c->ev_nn0) (void)hipEventDestroy(c->ev_nn0);
    if (c->ev_nn1) (void)hipEventDestroy(c->ev_nn1);
    for (hipEvent_t e : c->sweep_timing.ev)
        if (e) (void)hipEventDestroy(e);
    for (void* p : c->allocations) (void)hipFree(p);
    if (c->sc_host) (void)hipHostFree(c->sc_host);
    if (c->pred.sc_host) (void)hipHostFree(c->pred.sc_host);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

static int check_params(const sg_params* p) {
    if (!p) return SG_ERR_ARG;
    if (!(p->rho_l > 0.0) || !(p->rho_g > 0.0) || !(p->rho_ref > 0.0)) {
        set_error("densities must be positive");
        return SG_ERR_ARG;
    }
    if (p->precision != SG_PREC_FP64 && p->precision != SG_PREC_MIXED) {
        set_error("precision must be SG_PREC_FP64 or SG_PREC_MIXED");
        return SG_ERR_ARG;
    }
    if (!(p->tol > 0.0)) {
        set_error("tol must be positive");
        return SG_ERR_ARG;
    }
    if (p->mg_coarse_krylov < 0 || p->mg_coarse_krylov > 64) {
        set_error("mg_coarse_krylov must be in [0, 64]");
        return SG_ERR_ARG;
    }
    if (p->mg_pre != 1 || p->mg_post != 1) {
        set_error("mg_pre and mg_post must be 1 (the cycle kernels fuse one sweep each)");
        return SG_ERR_ARG;
    }
    if (p->warm_start < 0 || p->warm_start > 3) {
        set_error("warm_start must be 0, 1, 2 or 3");
        return SG_ERR_ARG;
    }
    if (p->momentum_method != 0 && p->momentum_method != 1) {
        set_error("momentum_method must be 0 (Jacobi) or 1 (BiCGStab)");
        return SG_ERR_ARG;
    }
    return SG_OK;
}

}  // namespace sg

using namespace sg;

extern "C" {

const char* sg_last_error(void) { return g_err.c_str(); }

int sg_params_default(sg_params* p) {
    if (!p) return SG_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    p->rho_l = 998.2;  // StateVariables.f90:27
    p->rho_g = 1.225;
    p->rho_ref = 998.2;  // Main.f90:148
    p->nu_l = 1.0034e-6;
    p->nu_g = 1.506e-5;
    p->nu_ref = 1.0034e-6;
    p->Re = 1.0;
    p->Fr = 1.0;
    p->bc_w = 1;
    p->bc_e = 1;
    p->bc_s = 1;
    p->bc_n = 0;
    p->precision = SG_PREC_FP64;
    p->max_iter = 500;
    p->tol = 1e-10;
    p->mg_pre = 1;
    p->mg_post = 1;
    p->warm_start = 0;
    p->check_every = 0;
    p->mg_coarse_krylov = 4;
    p->momentum_method = 0;
    return SG_OK;
}

int sg_ctx_create(int ni, int nj, const double* dx, const double* dy, const double* udx, const double* vdy,
                  const sg_params* prm, sg_ctx** out) {
    if (!out || ni < 2 || nj < 2 || !dx || !dy || !udx || !vdy) {
        set_error("sg_ctx_create: need ni, nj >= 2 and all spacing arrays");
        return SG_ERR_ARG;
    }
    *out = nullptr;
    int rc = check_params(prm);
    if (rc) return rc;
    if ((long long)ni * (long long)nj > (1ll << 31)) {
        set_error("sg_ctx_create: grid too large");
        return SG_ERR_ARG;
    }
    Ctx* c = new Ctx();
    if (const char* ev = std::getenv("SG_NO_GRAPH")) c->use_graph = (ev[0] == '0');
    c->ni = ni;
    c->nj = nj;
    c->prm = *prm;
#define CK(call)                    \
    do {                            \
        hipError_t _e = (call);     \
        if (_e != hipSuccess) {     \
            hip_fail(_e, #call);    \
            destroy(c);             \
            return SG_ERR_HIP;      \
        }                           \
    } while (0)
    CK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CK(hipEventCreate(&c->ev0));
    CK(hipEventCreate(&c->ev1));
    CK(hipEventCreate(&c->ev2));
    auto fail_alloc = [&]() {
        set_error("sg_ctx_create: device allocation failed");
        destroy(c);
        return SG_ERR_ALLOC;
    };
    c->dx = (double*)dev_alloc(c, ni * sizeof(double));
    c->dy = (double*)dev_alloc(c, nj * sizeof(double));
    c->udx = (double*)dev_alloc(c, ni * sizeof(double));
    c->vdy = (double*)dev_alloc(c, nj * sizeof(double));
    if (!c->dx || !c->dy || !c->udx || !c->vdy) return fail_alloc();
    CK(hipMemcpy(c->dx, dx, ni * sizeof(double), hipMemcpyHostToDevice));
    CK(hipMemcpy(c->dy, dy, nj * sizeof(double), hipMemcpyHostToDevice));
    CK(hipMemcpy(c->udx, udx, ni * sizeof(double), hipMemcpyHostToDevice));
    CK(hipMemcpy(c->vdy, vdy, nj * sizeof(double), hipMemcpyHostToDevice));
    const size_t ncell = (size_t)ni * nj, nghost = (size_t)(ni + 2) * (nj + 2);
    for (int f = 0; f < SG_F_COUNT; ++f) {
        size_t bytes = 0;
        if (is_cell_field(f)) bytes = ncell * sizeof(double);
        else if (is_ghost_field(f)) bytes = nghost * sizeof(double);
        else if (is_int_field(f)) bytes = ncell * sizeof(int);
        c->field_bytes[f] = bytes;
        c->field[f] = dev_alloc(c, bytes);
        if (!c->field[f]) return fail_alloc();
    }
    {
        double dxmin = dx[0], dymax = dy[0];
        for (int i = 1; i < ni; ++i) dxmin = std::min(dxmin, dx[i]);
        for (int j = 1; j < nj; ++j) dymax = std::max(dymax, dy[j]);
        c->semi0 = (nj >= 4 && dxmin >= SEMI_ASPECT * dymax) ? 1 : 0;
        if (const char* ev = std::getenv("SG_MG_SEMI")) c->semi0 = (ev[0] == '1') ? c->semi0 : 0;
    }
    if ((rc = mg_allocate(c)) != SG_OK) return fail_alloc();
    const size_t n0 = c->lev[0].d.alloc_elems();
    double** vecs[] = {&c->x, &c->r, &c->z, &c->p0, &c->p1, &c->q, &c->rhs};
    for (double** v : vecs) {
        *v = (double*)dev_alloc(c, n0 * sizeof(double));
        if (!*v) return fail_alloc();
    }
    float** vecf[] = {&c->rf, &c->zf, &c->pf0, &c->pf1, &c->qf};
    for (float** v : vecf) {
        *v = (float*)dev_alloc(c, n0 * sizeof(float));
        if (!*v) return fail_alloc();
    }
    if (clsvof_allocate(c) != SG_OK) return fail_alloc();
    if (predictor_allocate(c) != SG_OK) return fail_alloc();
    {
        // default cell centres: faces from 0 by the spacings (Mesh_Mod.f90:123-135)
        std::vector<double> xc(ni), yc(nj);
        double f = 0.0;
        for (int i = 0; i < ni; ++i) {
            xc[i] = 0.5 * ((f + dx[i]) + f);
            f = f + dx[i];
        }
        f = 0.0;
        for (int j = 0; j < nj; ++j) {
            yc[j] = 0.5 * ((f + dy[j]) + f);
            f = f + dy[j];
        }
        CK(hipMemcpy(c->cls.xc, xc.data(), ni * sizeof(double), hipMemcpyHostToDevice));
        CK(hipMemcpy(c->cls.yc, yc.data(), nj * sizeof(double), hipMemcpyHostToDevice));
        c->cls.xw = xc[0] - dx[0] / 2.0;
        c->cls.dx0 = dx[0];
    }
    // cell-layout staging for the host-pointer test hooks (5 planes)
    c->stage = (double*)dev_alloc(c, 5 * ncell * sizeof(double));
    if (!c->stage) return fail_alloc();
    const dim3 g = grid2d(ni, nj);
    c->red.max_blocks = (int)(g.x * g.y);
    c->red.partials = (double*)dev_alloc(c, sizeof(double) * 8 * NRED * (size_t)c->red.max_blocks);
    c->red.tickets = (unsigned int*)dev_alloc(c, sizeof(unsigned int) * 8);
    c->sc = (PcgScalars*)dev_alloc(c, sizeof(PcgScalars));
    c->step_dev = (double*)dev_alloc(c, 4 * sizeof(double));
    if (!c->red.partials || !c->red.tickets || !c->sc || !c->step_dev) return fail_alloc();
    c->diag.ring.rows = (double*)dev_alloc(c, sizeof(double) * SG_DIAG_RING * SG_DIAG_COLS);
    c->diag.ring.count = (long long*)dev_alloc(c, sizeof(long long));
    c->diag.partials = (double*)dev_alloc(c, sizeof(double) * 7 * DIAG_BLOCKS);
    if (!c->diag.ring.rows || !c->diag.ring.count || !c->diag.partials) return fail_alloc();
    CK(hipHostMalloc((void**)&c->sc_host, sizeof(PcgScalars), hipHostMallocDefault));
    std::memset(c->sc_host, 0, sizeof(PcgScalars));
    // a geometry with no solid: open apertures, fluid everywhere
    {
        std::vector<double> ones(nghost, 1.0);
        for (int f : {SG_F_AE, SG_F_AW, SG_F_AN, SG_F_AS}) CK(hipMemcpy(c->field[f], ones.data(), nghost * sizeof(double), hipMemcpyHostToDevice));
        std::vector<int> flags(ncell, 1);
        for (int f : {SG_F_UFLAG, SG_F_VFLAG}) CK(hipMemcpy(c->field[f], flags.data(), ncell * sizeof(int), hipMemcpyHostToDevice));
    }
    CK(hipDeviceSynchronize());
#undef CK
    *out = reinterpret_cast<sg_ctx*>(c);
    return SG_OK;
}

int sg_ctx_destroy(sg_ctx* ctx) {
    destroy(reinterpret_cast<Ctx*>(ctx));
    return SG_OK;
}

int sg_set_params(sg_ctx* ctx, const sg_params* p) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    int rc = check_params(p);
    if (rc) return rc;
    const bool prec_changed = p->precision != c->prm.precision;
    const bool coef_changed = p->rho_l != c->prm.rho_l || p->rho_g != c->prm.rho_g || p->rho_ref != c->prm.rho_ref ||
                              p->bc_w != c->prm.bc_w || p->bc_e != c->prm.bc_e || p->bc_s != c->prm.bc_s ||
                              p->bc_n != c->prm.bc_n;
    c->prm = *p;
    if (prec_changed || coef_changed) {
        c->built = 0;
        c->mg_ready = 0;
    }
    return SG_OK;
}

void* sg_field_ptr(sg_ctx* ctx, int field) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || field < 0 || field >= SG_F_COUNT) return nullptr;
    return c->field[field];
}

int sg_upload(sg_ctx* ctx, int field, const void* host) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !host || field < 0 || field >= SG_F_COUNT) {
        set_error("sg_upload: bad argument");
        return SG_ERR_ARG;
    }
    SG_HIP(hipMemcpyAsync(c->field[field], host, c->field_bytes[field], hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    c->built = 0;
    return SG_OK;
}

int sg_download(sg_ctx* ctx, int field, void* host) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !host || field < 0 || field >= SG_F_COUNT) {
        set_error("sg_download: bad argument");
        return SG_ERR_ARG;
    }
    SG_HIP(hipMemcpyAsync(host, c->field[field], c->field_bytes[field], hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

// device-to-device variants: the other buffer lives on the same GPU (a
// caller's own allocation); synchronous on return
int sg_upload_device(sg_ctx* ctx, int field, const void* dev) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !dev || field < 0 || field >= SG_F_COUNT) {
        set_error("sg_upload_device: bad argument");
        return SG_ERR_ARG;
    }
    SG_HIP(hipMemcpyAsync(c->field[field], dev, c->field_bytes[field], hipMemcpyDeviceToDevice, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    c->built = 0;
    return SG_OK;
}

int sg_download_device(sg_ctx* ctx, int field, void* dev) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !dev || field < 0 || field >= SG_F_COUNT) {
        set_error("sg_download_device: bad argument");
        return SG_ERR_ARG;
    }
    SG_HIP(hipMemcpyAsync(dev, c->field[field], c->field_bytes[field], hipMemcpyDeviceToDevice, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_synchronize(sg_ctx* ctx) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_set_geometry(sg_ctx* ctx, const double* vofS, const double* nyS, const double* WlLh, const double* aE,
                    const double* aW, const double* aN, const double* aS, const int* uflag, const int* vflag,
                    const double* vofSU, const double* vofSV) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !vofS || !nyS || !WlLh || !aE || !aW || !aN || !aS || !uflag || !vflag || !vofSU || !vofSV) {
        set_error("sg_set_geometry: null argument");
        return SG_ERR_ARG;
    }
    const void* src[] = {vofS, nyS, WlLh, aE, aW, aN, aS, uflag, vflag, vofSU, vofSV};
    const int dst[] = {SG_F_VOFS, SG_F_NYS, SG_F_WLLH, SG_F_AE, SG_F_AW, SG_F_AN, SG_F_AS,
                       SG_F_UFLAG, SG_F_VFLAG, SG_F_VOFSU, SG_F_VOFSV};
    for (int k = 0; k < 11; ++k)
        SG_HIP(hipMemcpyAsync(c->field[dst[k]], src[k], c->field_bytes[dst[k]], hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    c->built = 0;
    c->mg_ready = 0;
    return SG_OK;
}

int sg_poisson_build_resident(sg_ctx* ctx, double vb) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    c->vb = vb;
    return poisson_build(c);
}

int sg_poisson_solve_resident(sg_ctx* ctx, sg_solve_info* info) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    if (!c->built) {
        set_error("sg_poisson_solve_resident: call sg_poisson_build_resident first");
        return SG_ERR_STATE;
    }
    // the build's value guard is read after the solve, which runs either
    // way (its own NaN test stops it early on bit 1), so no wait comes
    // between the build and the solve.  A solve that returns SG_OK has
    // ended with the true-residual read-back, which copied every scalar,
    // the guard included, to the host mirror.
    int rc = pcg_solve(c, info);
    const int guard = rc == SG_OK ? c->sc_host->guard : 0;
    // clear the guard bits for the next build (stream-ordered)
    SG_HIP(hipMemsetAsync(&c->sc->guard, 0, sizeof(int), c->stream));
    if (guard && rc == SG_OK) {
        rc = (guard & 1) ? SG_NAN : SG_GUARD;
        if (info) info->status = rc;
        set_error("projection: NaN or out-of-range matrix/RHS value (ProjectionP_Mod.f90:209-213,322-325)");
    }
    if (rc == SG_NOT_CONVERGED) set_error("projection: PCG did not reach the tolerance within max_iter");
    return rc;
}

int sg_poisson_solve(sg_ctx* ctx, const double* u_star, const double* v_star, const double* DpU, const double* DpV,
                     const double* vof, const double* phi, double vb, double* Pp_out, sg_solve_info* info) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !u_star || !v_star || !DpU || !DpV || !vof || !phi || !Pp_out) {
        set_error("sg_poisson_solve: null argument");
        return SG_ERR_ARG;
    }
    const void* src[] = {u_star, v_star, DpU, DpV, vof, phi};
    const int dst[] = {SG_F_UPRED, SG_F_VPRED, SG_F_DPU, SG_F_DPV, SG_F_VOF, SG_F_PHI};
    for (int k = 0; k < 6; ++k)
        SG_HIP(hipMemcpyAsync(c->field[dst[k]], src[k], c->field_bytes[dst[k]], hipMemcpyHostToDevice, c->stream));
    int rc = sg_poisson_build_resident(ctx, vb);
    if (rc) return rc;
    rc = sg_poisson_solve_resident(ctx, info);
    if (rc < 0) return rc;
    SG_HIP(hipMemcpyAsync(Pp_out, c->field[SG_F_PP], c->field_bytes[SG_F_PP], hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return rc;
}

int sg_poisson_matrix(sg_ctx* ctx, double* matr_out) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !matr_out) return SG_ERR_ARG;
    const size_t ncell = (size_t)c->ni * c->nj;
    int rc = poisson_matrix_rows(c, c->stage);
    if (rc) return rc;
    SG_HIP(hipMemcpyAsync(matr_out, c->stage, 5 * ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_poisson_rhs(sg_ctx* ctx, double* rhs_out) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !rhs_out) return SG_ERR_ARG;
    if (!c->built) {
        set_error("sg_poisson_rhs: call sg_poisson_build_resident first");
        return SG_ERR_STATE;
    }
    int rc = poisson_rhs_cells(c, c->stage);
    if (rc) return rc;
    SG_HIP(hipMemcpyAsync(rhs_out, c->stage, (size_t)c->ni * c->nj * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_poisson_apply(sg_ctx* ctx, const double* x, double* y) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !x || !y) return SG_ERR_ARG;
    if (!c->built) {
        set_error("sg_poisson_apply: call sg_poisson_build_resident first");
        return SG_ERR_STATE;
    }
    const size_t ncell = (size_t)c->ni * c->nj;
    // padded staging vectors: level-0 x and r (interior writes only)
    double* xp = c->lev[0].v.x;
    double* yp = c->lev[0].v.r;
    SG_HIP(hipMemcpyAsync(c->stage, x, ncell * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = cell_to_padded(c, c->stage, xp);
    if (rc) return rc;
    rc = poisson_apply_fine(c, xp, yp);
    if (rc) return rc;
    rc = padded_to_cell(c, yp, c->stage);
    if (rc) return rc;
    SG_HIP(hipMemcpyAsync(y, c->stage, ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_velocity_correct_resident(sg_ctx* ctx, double vb) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    c->vb = vb;
    int rc = velocity_correct(c);
    if (rc) return rc;
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int sg_set_staggered_geometry(sg_ctx* ctx, int which, const sg_stag_geometry* g) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !g || (which != 0 && which != 1)) {
        set_error("sg_set_staggered_geometry: bad argument");
        return SG_ERR_ARG;
    }
    const double* src[STG_COUNT] = {g->vofS, g->nxS, g->nyS, g->phiS, g->WlLh, g->delh, g->ccx, g->ccy,
                                    g->aE, g->aW, g->aN, g->aS, g->EtaE, g->EtaN, g->AlE, g->AlN,
                                    g->DAlE, g->DAlN, g->SxE, g->SyN, g->fcex, g->fcey, g->fcnx, g->fcny};
    for (int k = 0; k < STG_COUNT; ++k)
        if (!src[k]) {
            set_error("sg_set_staggered_geometry: null array");
            return SG_ERR_ARG;
        }
    if (!g->moex || !g->active) {
        set_error("sg_set_staggered_geometry: null array");
        return SG_ERR_ARG;
    }
    const size_t n = (size_t)c->ni * c->nj, ng = (size_t)(c->ni + 2) * (c->nj + 2);
    StagDev& S = c->stag[which];
    for (int k = 0; k < STG_COUNT; ++k)
        SG_HIP(hipMemcpyAsync(S.a[k], src[k], (k < STG_NCELL ? n : ng) * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
    SG_HIP(hipMemcpyAsync(S.moex, g->moex, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    // the reference numbers U cells of columns 1..ni-1 and V cells of rows
    // 1..nj-1 only (NumberExternalCell, Cutcell_Mod.f90:340-358)
    std::vector<int> act(g->active, g->active + n);
    for (int j = 0; j < c->nj; ++j)
        for (int i = 0; i < c->ni; ++i)
            if ((which == 0 && i == c->ni - 1) || (which == 1 && j == c->nj - 1)) act[(size_t)i + (size_t)j * c->ni] = 0;
    SG_HIP(hipMemcpyAsync(S.act, act.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_predictor_resident(sg_ctx* ctx, double dt, double vb, int first_step, sg_predictor_info* info) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    if (!(dt > 0.0)) {
        set_error("sg_predictor_resident: dt must be positive");
        return SG_ERR_ARG;
    }
    if (c->ni < 4 || c->nj < 4) {
        set_error("sg_predictor_resident: needs ni, nj >= 4");
        return SG_ERR_ARG;
    }
    SG_HIP(hipEventRecord(c->ev0, c->stream));
    int rc = predictor_assemble(c, dt, vb, first_step);
    if (rc) return rc;
    sg_predictor_info loc{};
    const double tol = 1.0e-13;
    rc = momentum_solve(c, c->pred.mu, c->pred.ru, c->pred.xu, tol, 500, &loc.iterations_u, &loc.relres_u);
    if (rc < 0) return rc;
    int rc2 = momentum_solve(c, c->pred.mv, c->pred.rv, c->pred.xv, tol, 500, &loc.iterations_v, &loc.relres_v);
    if (rc2 < 0) return rc2;
    // the time is taken up to the end of the solves, whose convergence the
    // host has just read; the scatter and Dp/velocity finish are enqueued
    // behind the event and not waited for
    SG_HIP(hipEventRecord(c->ev1, c->stream));
    int rc3 = predictor_finish(c);
    if (rc3) return rc3;
    SG_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    SG_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    loc.ms = ms;
    if (info) *info = loc;
    c->built = 0;
    if (rc == SG_NAN || rc2 == SG_NAN) return SG_NAN;  // the solve's message stands
    if (rc || rc2) {
        set_error("sg_predictor_resident: a momentum system did not converge");
        return SG_NOT_CONVERGED;
    }
    return SG_OK;
}

int sg_predictor_rows(sg_ctx* ctx, double* matr_u, double* rhs_u, double* matr_v, double* rhs_v) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !matr_u || !rhs_u || !matr_v || !rhs_v) return SG_ERR_ARG;
    const size_t n = (size_t)c->ni * c->nj;
    SG_HIP(hipMemcpyAsync(matr_u, c->pred.mu, 5 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipMemcpyAsync(rhs_u, c->pred.ru, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipMemcpyAsync(matr_v, c->pred.mv, 5 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipMemcpyAsync(rhs_v, c->pred.rv, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_particles_set(sg_ctx* ctx, int n, const double* x, const double* y, const double* u, const double* v,
                     const double* d) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || n < 0 || (n > 0 && (!x || !y || !u || !v || !d))) {
        set_error("sg_particles_set: need n >= 0 and five host arrays of n values");
        return SG_ERR_ARG;
    }
    ParticleState& S = c->part;
    const int rc = particles_reserve(c, n);
    if (rc) return rc;
    S.n = n;
    if (n == 0) return SG_OK;
    const size_t b = (size_t)n * sizeof(double);
    SG_HIP(hipMemcpyAsync(S.x, x, b, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(S.y, y, b, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(S.u, u, b, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(S.v, v, b, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(S.d, d, b, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_particles_get(sg_ctx* ctx, int n, double* x, double* y, double* u, double* v) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || n != c->part.n || (n > 0 && (!x || !y || !u || !v))) {
        set_error("sg_particles_get: n must be the count given to sg_particles_set, with four host arrays");
        return SG_ERR_ARG;
    }
    if (n == 0) return SG_OK;
    const ParticleState& S = c->part;
    const size_t b = (size_t)n * sizeof(double);
    SG_HIP(hipMemcpyAsync(x, S.x, b, hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipMemcpyAsync(y, S.y, b, hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipMemcpyAsync(u, S.u, b, hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipMemcpyAsync(v, S.v, b, hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

// the checks a solid shape passes before any kernel draws it: a known kind,
// a finite centre, positive sizes, and for a boom a bar narrower than its
// float (b < a) of non-negative length; who names the caller in the error
static bool validate_shape(const sg_solid_shape* shape, const char* who) {
    if (!shape || shape->kind < SG_SOLID_NONE || shape->kind > SG_SOLID_BOOM) {
        set_error(std::string(who) + ": need a shape of a known kind");
        return false;
    }
    if (shape->kind != SG_SOLID_NONE &&
        (!(shape->a > 0.0) || (shape->kind != SG_SOLID_CIRCLE && !(shape->b > 0.0)) ||
         !std::isfinite(shape->cx) || !std::isfinite(shape->cy))) {
        set_error(std::string(who) + ": the shape needs a finite centre and positive sizes");
        return false;
    }
    if (shape->kind == SG_SOLID_BOOM && (!(shape->b < shape->a) || !(shape->c >= 0.0))) {
        set_error(std::string(who) + ": a boom needs bar half width b < radius a and bar length c >= 0");
        return false;
    }
    return true;
}

int sg_geometry_rebuild(sg_ctx* ctx, const sg_solid_shape* shape) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!validate_shape(shape, "sg_geometry_rebuild")) return SG_ERR_ARG;
    if (!c) {
        set_error("sg_geometry_rebuild: null context");
        return SG_ERR_ARG;
    }
    if (c->ni < 4 || c->nj < 4) {
        set_error("sg_geometry_rebuild: needs ni, nj >= 4");
        return SG_ERR_ARG;
    }
    return geometry_rebuild(c, shape);
}

int sg_object_force(sg_ctx* ctx, double vs, double* force, double* clp, double* clf) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !force) {
        set_error("sg_object_force: null argument");
        return SG_ERR_ARG;
    }
    return object_force(c, vs, force, clp, clf);
}

int sg_set_moving_solid(sg_ctx* ctx, const sg_solid_shape* shape, double us, double vs, double asx, double asy) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!validate_shape(shape, "sg_set_moving_solid")) return SG_ERR_ARG;
    if (!std::isfinite(us) || !std::isfinite(vs) || !std::isfinite(asx) || !std::isfinite(asy)) {
        set_error("sg_set_moving_solid: velocities and accelerations must be finite");
        return SG_ERR_ARG;
    }
    if (!c) {
        set_error("sg_set_moving_solid: null context");
        return SG_ERR_ARG;
    }
    c->moving.shape = *shape;
    c->moving.us = us;
    c->moving.vs = vs;
    c->moving.asx = asx;
    c->moving.asy = asy;
    return SG_OK;
}

int sg_get_moving_solid(sg_ctx* ctx, sg_solid_shape* shape, double* us, double* vs) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !shape) {
        set_error("sg_get_moving_solid: null argument");
        return SG_ERR_ARG;
    }
    *shape = c->moving.shape;
    if (us) *us = c->moving.us;
    if (vs) *vs = c->moving.vs;
    return SG_OK;
}

int sg_geometry_snapshot(sg_ctx* ctx) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    return geometry_snapshot(c);
}

int sg_inter_new_var(sg_ctx* ctx, double vb) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    return inter_new_var(c, vb);
}

int sg_stag_download(sg_ctx* ctx, int which, int k, void* host) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !host || (which != 0 && which != 1) || k < 0 || k > STG_COUNT + 1) {
        set_error("sg_stag_download: bad argument");
        return SG_ERR_ARG;
    }
    const size_t n = (size_t)c->ni * c->nj, ng = (size_t)(c->ni + 2) * (c->nj + 2);
    const StagDev& S = c->stag[which];
    const void* src = k < STG_COUNT ? (const void*)S.a[k] : (k == STG_COUNT ? (const void*)S.moex : (const void*)S.act);
    const size_t bytes = k < STG_NCELL ? n * sizeof(double) : (k < STG_COUNT ? ng * sizeof(double) : n * sizeof(int));
    SG_HIP(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int sg_diagnostics_record(sg_ctx* ctx) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    return diagnostics_record(c);
}

int sg_diagnostics_read(sg_ctx* ctx, int max_rows, double* rows, int* n_rows, long long* n_lost) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || max_rows < 0 || (max_rows > 0 && !rows) || !n_rows) {
        set_error("sg_diagnostics_read: need max_rows >= 0, a row buffer and n_rows");
        return SG_ERR_ARG;
    }
    return diagnostics_read(c, max_rows, rows, n_rows, n_lost);
}

int sg_particles_track(sg_ctx* ctx, double dt, const sg_particle_params* p, int* flagged) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !p || !(p->lref > 0.0) || !(p->uref > 0.0)) {
        set_error("sg_particles_track: need parameters with lref, uref > 0");
        return SG_ERR_ARG;
    }
    return particles_track(c, dt, p, flagged);
}

int sg_interface_to_uv_resident(sg_ctx* ctx) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    if (c->ni < 4 || c->nj < 4) {
        set_error("sg_interface_to_uv_resident: needs ni, nj >= 4");
        return SG_ERR_ARG;
    }
    int rc = interface_to_uv(c);
    if (rc) return rc;
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int sg_clsvof_params_default(sg_clsvof_params* p) {
    if (!p) return SG_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    p->nsub = 6;          // Clsvof_Mod.f90:754
    p->depth = 1.0e300;   // inflow open on every row
    return SG_OK;
}

int sg_set_centres(sg_ctx* ctx, const double* xc, const double* yc) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !xc || !yc) {
        set_error("sg_set_centres: null argument");
        return SG_ERR_ARG;
    }
    SG_HIP(hipMemcpyAsync(c->cls.xc, xc, c->ni * sizeof(double), hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipMemcpyAsync(c->cls.yc, yc, c->nj * sizeof(double), hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    c->cls.xw = xc[0] - c->cls.dx0 / 2.0;
    return SG_OK;
}

int sg_clsvof_advect_resident(sg_ctx* ctx, double dt, double time, const sg_clsvof_params* p) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !p) return SG_ERR_ARG;
    if (c->ni < 4 || c->nj < 4) {
        set_error("sg_clsvof_advect: the transport needs ni, nj >= 4");
        return SG_ERR_ARG;
    }
    if (p->nsub < 1 || p->nsub > 1024) {
        set_error("sg_clsvof_advect: nsub must be in [1, 1024]");
        return SG_ERR_ARG;
    }
    if (!(dt >= 0.0)) {
        set_error("sg_clsvof_advect: dt must be >= 0");
        return SG_ERR_ARG;
    }
    // Depthw + etau of each sub-step (X_Sweep, Clsvof_Mod.f90:1057-1059)
    std::vector<double> levels(p->nsub);
    const double dtv = dt / (double)p->nsub;
    for (int k = 1; k <= p->nsub; ++k) {
        const double etau = p->amp * std::sin(p->kw * c->cls.xw - p->omega * (time + (double)k * dtv));
        levels[k - 1] = p->depth + etau;
    }
    int rc = clsvof_advect(c, dt, p->nsub, levels.data());
    if (rc) return rc;
    SG_HIP(hipGetLastError());
    c->built = 0;
    return SG_OK;
}

int sg_clsvof_advect(sg_ctx* ctx, double* vof, double* phi, double* nx, double* ny, const double* u, const double* v,
                     double dt, double time, const sg_clsvof_params* p) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !vof || !phi || !u || !v || !p) {
        set_error("sg_clsvof_advect: null argument");
        return SG_ERR_ARG;
    }
    const void* src[] = {vof, phi, u, v};
    const int dst[] = {SG_F_VOF, SG_F_PHI, SG_F_U, SG_F_V};
    for (int k = 0; k < 4; ++k)
        SG_HIP(hipMemcpyAsync(c->field[dst[k]], src[k], c->field_bytes[dst[k]], hipMemcpyHostToDevice, c->stream));
    int rc = sg_clsvof_advect_resident(ctx, dt, time, p);
    if (rc) return rc;
    SG_HIP(hipMemcpyAsync(vof, c->field[SG_F_VOF], c->field_bytes[SG_F_VOF], hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipMemcpyAsync(phi, c->field[SG_F_PHI], c->field_bytes[SG_F_PHI], hipMemcpyDeviceToHost, c->stream));
    if (nx) SG_HIP(hipMemcpyAsync(nx, c->field[SG_F_NX], c->field_bytes[SG_F_NX], hipMemcpyDeviceToHost, c->stream));
    if (ny) SG_HIP(hipMemcpyAsync(ny, c->field[SG_F_NY], c->field_bytes[SG_F_NY], hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

}  // extern "C"

// one untimed launch, then reps launches between two events on the context's stream
template <class L>
static int time_launches(sg::Ctx* c, int reps, double* avg_ms, L launch) {
    int rc = launch();
    if (rc) return rc;
    SG_HIP(hipEventRecord(c->ev0, c->stream));
    for (int k = 0; k < reps; ++k) {
        rc = launch();
        if (rc) return rc;
    }
    SG_HIP(hipEventRecord(c->ev1, c->stream));
    SG_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    SG_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = (double)ms / reps;
    return SG_OK;
}

extern "C" {

int sg_sweep_timing(sg_ctx* ctx, int on, long long* launches, double* total_ms) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    auto& T = c->sweep_timing;
    if (on) {
        if (!T.ev[0])
            for (hipEvent_t& e : T.ev) SG_HIP(hipEventCreate(&e));
        T.on = 1;
        T.used = 0;
        T.launches = 0;
        T.ms = 0.0;
        return SG_OK;
    }
    int rc = T.on ? sweep_timing_drain(c) : SG_OK;
    T.on = 0;
    if (launches) *launches = T.launches;
    if (total_ms) *total_ms = T.ms;
    return rc;
}

int sg_time_kernel(sg_ctx* ctx, int which, int reps, double* avg_ms) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !avg_ms || reps <= 0) return SG_ERR_ARG;
    if (which == 4) {
        // the particle tracker on the current set, with the last call's dt
        // and parameters (the set advances reps + 1 steps)
        if (!c->part.tracked) {
            set_error("sg_time_kernel: track the particles once before timing them");
            return SG_ERR_STATE;
        }
        return time_launches(c, reps, avg_ms, [&]() { return particles_launch(c, c->part.last_dt, &c->part.last); });
    }
    if (which == 5) return time_launches(c, reps, avg_ms, [&]() { return diagnostics_record(c); });
    if (!c->built) {
        set_error("sg_time_kernel: build the operator first");
        return SG_ERR_STATE;
    }
    if (which < 0 || which > 5) {
        set_error("sg_time_kernel: unknown kernel id");
        return SG_ERR_ARG;
    }
    if (which > 0 && !c->mg_ready) {
        set_error("sg_time_kernel: solve once before timing the cycle kernels");
        return SG_ERR_STATE;
    }
    // 0: fine operator apply y = A x (PCG x -> level-0 residual buffer);
    // 1/2: fine down/up sweep of the V-cycle; 3: plain copy (calibration)
    return time_launches(c, reps, avg_ms, [&]() {
        return which == 0 ? poisson_apply_fine(c, c->x, c->lev[0].v.r) : mg_launch_fine(c, which);
    });
}

}  // extern "C"

// ---- learned warm start ---------------------------------------------------
struct sg_nn {
    sg::nn::Net net;
    hipStream_t stream = nullptr;
    float* host_loss = nullptr;
};

extern "C" {

size_t sg_nn_num_weights(void) { return sg::nn::num_weights(); }

int sg_nn_init_weights(float* w, unsigned long long seed) {
    using namespace sg::nn;
    if (!w) return SG_ERR_ARG;
    // xorshift64* + Box-Muller; He-normal std sqrt(2 / (9 cin)) on the first
    // 9*cin columns of each row, zero padding; the output layer starts small
    // and only its row 0 is nonzero
    unsigned long long st = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
    auto uni = [&]() {
        st ^= st >> 12;
        st ^= st << 25;
        st ^= st >> 27;
        return ((st * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
    };
    size_t o = 0;
    for (int l = 0; l < NL; ++l) {
        const int cin = SPEC[l].cin, cout = SPEC[l].cout, kp = kpad(cin);
        const double sd = (l == NL - 1 ? 0.1 : 1.0) * std::sqrt(2.0 / (9.0 * cin));
        for (int co = 0; co < cout; ++co)
            for (int k = 0; k < kp; ++k) {
                double v = 0.0;
                if (k < 9 * cin && (l < NL - 1 || co == 0)) {
                    const double u1 = uni() + 1e-300, u2 = uni();
                    v = sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                }
                w[o++] = (float)v;
            }
    }
    return SG_OK;
}

int sg_nn_create(int H, int W, int batch, sg_nn** out) {
    if (!out) return SG_ERR_ARG;
    *out = nullptr;
    sg_nn* n = new sg_nn();
    if (hipStreamCreate(&n->stream) != hipSuccess || hipHostMalloc(&n->host_loss, sizeof(float)) != hipSuccess) {
        if (n->stream) (void)hipStreamDestroy(n->stream);
        delete n;
        sg::set_error("sg_nn_create: HIP stream / pinned memory");
        return SG_ERR_HIP;
    }
    int rc = sg::nn::create(&n->net, H, W, batch, n->stream);
    if (rc) {
        (void)hipStreamDestroy(n->stream);
        (void)hipHostFree(n->host_loss);
        delete n;
        return rc;
    }
    *out = n;
    return SG_OK;
}

int sg_nn_destroy(sg_nn* n) {
    if (!n) return SG_OK;
    (void)hipStreamSynchronize(n->stream);
    sg::nn::destroy(&n->net);
    (void)hipHostFree(n->host_loss);
    (void)hipStreamDestroy(n->stream);
    delete n;
    return SG_OK;
}

int sg_nn_set_weights(sg_nn* n, const float* w) {
    if (!n || !w) return SG_ERR_ARG;
    return sg::nn::set_weights(&n->net, w);
}

int sg_nn_get_weights(sg_nn* n, float* w) {
    if (!n || !w) return SG_ERR_ARG;
    return sg::nn::get_weights(&n->net, w);
}

int sg_nn_forward(sg_nn* n, const void* x, float* y, int nb) {
    if (!n || !x || !y) return SG_ERR_ARG;
    int rc = sg::nn::forward(&n->net, (const sg::nn::bf16*)x, nb);
    if (rc) return rc;
    SG_HIP(hipMemcpyAsync(y, n->net.y, (size_t)nb * n->net.H * n->net.W * sizeof(float), hipMemcpyDeviceToDevice,
                          n->stream));
    SG_HIP(hipStreamSynchronize(n->stream));
    return SG_OK;
}

static int nn_loss_grad(sg_nn* n, const void* x, const float* t, const float* mask, int nb, float* loss) {
    if (!n || !x || !t || !mask) return SG_ERR_ARG;
    int rc = sg::nn::backward(&n->net, (const sg::nn::bf16*)x, t, mask, nb);
    if (rc) return rc;
    if (loss) {
        SG_HIP(hipMemcpyAsync(n->host_loss, n->net.loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    }
    return SG_OK;
}

int sg_nn_gradient(sg_nn* n, const void* x, const float* t, const float* mask, int nb, float* grad, float* loss) {
    int rc = nn_loss_grad(n, x, t, mask, nb, loss);
    if (rc) return rc;
    if (grad) SG_HIP(hipMemcpyAsync(grad, n->net.g, n->net.nw * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    SG_HIP(hipStreamSynchronize(n->stream));
    if (loss) *loss = *n->host_loss;
    return SG_OK;
}

int sg_nn_train_step(sg_nn* n, const void* x, const float* t, const float* mask, int nb, float lr, float* loss) {
    int rc = nn_loss_grad(n, x, t, mask, nb, loss);
    if (rc) return rc;
    rc = sg::nn::adam_step(&n->net, lr);
    if (rc) return rc;
    if (loss) {
        SG_HIP(hipStreamSynchronize(n->stream));
        *loss = *n->host_loss;
    }
    return SG_OK;
}

int sg_nn_level_dims(sg_ctx* ctx, int level, int* H, int* W) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !H || !W || level < 0 || level >= c->nlev) return SG_ERR_ARG;
    *H = c->lev[level].d.nj;
    *W = c->lev[level].d.ni;
    return SG_OK;
}

int sg_nn_features(sg_ctx* ctx, int level, void* x, float* t, float* mask, double* scale) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !x || (t && !mask)) return SG_ERR_ARG;
    return sg::learned_features(c, level, x, t, mask, scale);
}

int sg_set_learned(sg_ctx* ctx, const float* w, int level, double out_scale, int deflate) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    int rc = sg::learned_enable(c, w, level, out_scale, deflate);
    if (rc == SG_OK) (void)hipStreamSynchronize(c->stream);
    return rc;
}

int sg_set_learned_prediction(sg_ctx* ctx, int level, const double* e, int deflate) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    return sg::learned_set_prediction(c, level, e, deflate);
}

}  // extern "C"

// ---- time-step control ------------------------------------------------------
extern "C" {

int sg_time_step_size(sg_ctx* ctx, double cfl, double vint, double uref, double vs, double asy, double* dt) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !dt || !(cfl > 0.0) || uref == 0.0) {
        sg::set_error("sg_time_step_size: bad argument");
        return SG_ERR_ARG;
    }
    return sg::time_step_size(c, cfl, vint, uref, vs, asy, dt);
}

int sg_time_step_size_noalloc(int ni, int nj, const double* udx, const double* vdy, const double* u, const double* v,
                              double cfl, double Fr, double vint, double uref, double vs, double asy, double* dt,
                              void* stream) {
    if (!udx || !vdy || !u || !v || !dt || !(cfl > 0.0) || uref == 0.0) {
        sg::set_error("sg_time_step_size_noalloc: bad argument");
        return SG_ERR_ARG;
    }
    return sg::time_step_size_noalloc(ni, nj, udx, vdy, u, v, cfl, Fr, vint, uref, vs, asy, dt,
                                      reinterpret_cast<hipStream_t>(stream));
}

int sg_closed_box_bc(sg_ctx* ctx) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return SG_ERR_ARG;
    int rc = sg::closed_box_bc(c);
    if (rc) return rc;
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int sg_field_copy(sg_ctx* ctx, int dst, int src) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || dst < 0 || src < 0 || dst >= SG_F_COUNT || src >= SG_F_COUNT ||
        c->field_bytes[dst] != c->field_bytes[src]) {
        sg::set_error("sg_field_copy: fields of different shape or out of range");
        return SG_ERR_ARG;
    }
    SG_HIP(hipMemcpyAsync(c->field[dst], c->field[src], c->field_bytes[src], hipMemcpyDeviceToDevice, c->stream));
    c->built = 0;
    return SG_OK;
}

}  // extern "C"
