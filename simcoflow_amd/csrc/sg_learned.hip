// Learned warm start and deflation vector of the projection PCG (no
// counterpart in the reference, which starts from x0 = 0,
// ProjectionP_Mod.f90:319).
//
// The net runs on level L of the multigrid hierarchy.  Its 8 input channels
// at coarse point (I,J): the right-hand side restricted L times with the
// cycle's own restriction P^T, divided by its rms over the fluid points; the
// four couplings of the Galerkin operator divided by its diagonal (density
// jumps and apertures); the liquid fraction averaged over the 2^L x 2^L
// fine block; and the point's position (y, x) in [0,1].  Its output times
// out_scale * rms is a pressure on level L; the cycle's prolongation carries
// it to the fine grid (w), PCG starts from x0 = (w,b)/(w,Aw) w and, with
// deflation on, keeps its directions A-orthogonal to w.
#include "sg_internal.hpp"
#include "sg_mg_common.hpp"
#include "sg_nn.hpp"
#include "sg_reduce.hpp"

namespace sg {

namespace {

using nn::bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int NTR = 1024;

// b_{l+1} = P^T b_l (semi: the level coarsens in y only, coarse (I,J) on
// fine (I,2J))
__global__ void __launch_bounds__(256) k_restrict(Dims F, Dims C, Prol<double> P, const double* __restrict__ bf,
                                                  double* __restrict__ bc, int semi) {
    const int I = blockIdx.x * TX + threadIdx.x;
    const int J = blockIdx.y * TY + threadIdx.y;
    if (I >= C.ni || J >= C.nj) return;
    const size_t c = C.off(I, J);
    if (semi) {
        bc[c] = P.w[W_INJ][c] * bf[F.off(I, 2 * J)] + P.w[W_YS][c] * bf[F.off(I, 2 * J + 1)] +
                P.w[W_YN][c - C.ld] * bf[F.off(I, 2 * J - 1)];
        return;
    }
    auto R = [&](int a, int b) { return bf[F.off(2 * I + a, 2 * J + b)]; };
    bc[c] = restrict_point(C, P, I, J, R);
}

// x_l = P e_{l+1}; solid points of the target level (diag = 0) get 0
__global__ void __launch_bounds__(256) k_prolong(Dims F, Dims C, Prol<double> P, const double* __restrict__ diag,
                                                 const double* __restrict__ e, double* __restrict__ x, int semi) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= F.ni || j >= F.nj) return;
    const size_t p = F.off(i, j);
    double v = 0.0;
    if (diag[p] > 0.0) {
        if (semi) {
            const size_t c = C.off(i, j >> 1);
            v = (j & 1) ? P.w[W_YS][c] * e[c] + P.w[W_YN][c] * e[c + C.ld] : P.w[W_INJ][c] * e[c];
        } else {
            v = prolong_point(C, P, e, i, j);
        }
    }
    x[p] = v;
}

// rms of b over the fluid points of the level -> sc->nn_scale, in two
// phases: per-block partials of (sum b^2, count), then one block
__global__ void __launch_bounds__(256) k_rms(Dims D, const double* __restrict__ diag, const double* __restrict__ b,
                                             double* __restrict__ part, int max_blocks) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    double v[2] = {0.0, 0.0};
    if (i < D.ni && j < D.nj) {
        const size_t p = D.off(i, j);
        if (diag[p] > 0.0) {
            v[0] = b[p] * b[p];
            v[1] = 1.0;
        }
    }
    block_partials<2, TX * TY>(v, part, max_blocks);
}

__global__ void __launch_bounds__(NTR) k_rms_fin(const double* __restrict__ part, int max_blocks, int nb,
                                                 PcgScalars* sc) {
    double s[2];
    finalize_sum<2, NTR>(part, max_blocks, nb, s);
    if (threadIdx.x == 0) {
        const double r = s[1] > 0.0 ? sqrt(s[0] / s[1]) : 0.0;
        sc->nn_scale = r > 0.0 ? r : 1.0;
    }
}

// sum and count of the fine cells of block (I,J), rows outer: all S*S loads
// are independent, so they are in flight together
template <int S>
__device__ inline double block_vof(const double* __restrict__ vof, int I, int J, int ni0, int nj0, int& nv) {
    double vs = 0.0;
    nv = 0;
#pragma unroll
    for (int a = 0; a < S; ++a)
#pragma unroll
        for (int c = 0; c < S; ++c) {
            const int jj = J * S + a, ii = I * S + c;
            if (jj < nj0 && ii < ni0) {
                vs += vof[(size_t)jj * ni0 + ii];
                ++nv;
            }
        }
    return vs;
}

// net input at level L (see the file comment); optionally the training
// target t = x(sx I, 2^L J) / scale and the fluid mask.  A level-L point
// covers sx x 2^L fine cells: sx = 2^L, or 2^(L-1) when level 1 coarsens in
// y only (semi)
__global__ void __launch_bounds__(256) k_features(Dims D, Op<double> A, const double* __restrict__ b, Dims F0,
                                                  const double* __restrict__ vof, int ni0, int nj0, int L, int semi,
                                                  const double* __restrict__ x0, const PcgScalars* sc,
                                                  bf16* __restrict__ X, float* __restrict__ t,
                                                  float* __restrict__ mask) {
    const int I = blockIdx.x * TX + threadIdx.x;
    const int J = blockIdx.y * TY + threadIdx.y;
    if (I >= D.ni || J >= D.nj) return;
    const size_t p = D.off(I, J);
    const double inv = 1.0 / sc->nn_scale;
    const double d = A.diag[p];
    const double id = d > 0.0 ? 1.0 / d : 0.0;
    const int s = 1 << L;
    const int sx = (semi && L > 0) ? s / 2 : s;
    double vs = 0.0;
    int nv = 0;
    switch (sx == s ? L : -1) {
    case 0: vs = block_vof<1>(vof, I, J, ni0, nj0, nv); break;
    case 1: vs = block_vof<2>(vof, I, J, ni0, nj0, nv); break;
    case 2: vs = block_vof<4>(vof, I, J, ni0, nj0, nv); break;
    case 3: vs = block_vof<8>(vof, I, J, ni0, nj0, nv); break;
    case 4: vs = block_vof<16>(vof, I, J, ni0, nj0, nv); break;
    default:
        for (int jj = J * s; jj < (J + 1) * s && jj < nj0; ++jj)
            for (int ii = I * sx; ii < (I + 1) * sx && ii < ni0; ++ii) {
                vs += vof[(size_t)jj * ni0 + ii];
                ++nv;
            }
    }
    float f[8];
    f[0] = d > 0.0 ? (float)(b[p] * inv) : 0.f;
    f[1] = (float)(A.cE[p] * id);
    f[2] = (float)(A.cN[p] * id);
    f[3] = A.cNE ? (float)(A.cNE[p] * id) : 0.f;
    f[4] = A.cNW ? (float)(A.cNW[p] * id) : 0.f;
    f[5] = nv ? (float)(vs / nv) : 0.f;
    f[6] = (float)((J + 0.5) / D.nj);
    f[7] = (float)((I + 0.5) / D.ni);
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = (bf16)f[k];
    *(bf16x8*)(X + ((size_t)(J + 1) * (D.ni + 2) + (I + 1)) * 8) = o;
    if (t) {
        const size_t q = (size_t)J * D.ni + I;
        t[q] = d > 0.0 ? (float)(x0[F0.off(I * sx, J * s)] * inv) : 0.f;
        mask[q] = d > 0.0 ? 1.f : 0.f;
    }
}

// level-L pressure from the net output
__global__ void __launch_bounds__(256) k_from_net(Dims D, const double* __restrict__ diag, const float* __restrict__ y,
                                                  double k, const PcgScalars* sc, double* __restrict__ e) {
    const int I = blockIdx.x * TX + threadIdx.x;
    const int J = blockIdx.y * TY + threadIdx.y;
    if (I >= D.ni || J >= D.nj) return;
    const size_t p = D.off(I, J);
    e[p] = diag[p] > 0.0 ? k * sc->nn_scale * (double)y[(size_t)J * D.ni + I] : 0.0;
}

// level-L prediction supplied by the caller; solid points of the level get 0
__global__ void __launch_bounds__(256) k_take_prediction(Dims D, const double* __restrict__ diag,
                                                         const double* __restrict__ ext, double* __restrict__ e) {
    const int I = blockIdx.x * TX + threadIdx.x;
    const int J = blockIdx.y * TY + threadIdx.y;
    if (I >= D.ni || J >= D.nj) return;
    const size_t p = D.off(I, J);
    e[p] = diag[p] > 0.0 ? ext[p] : 0.0;
}

// q = A w; partials of (w,b), (w,q)
__global__ void __launch_bounds__(256) k_galerkin(Dims D, const double* __restrict__ diag,
                                                  const double* __restrict__ cE, const double* __restrict__ cN,
                                                  const double* __restrict__ w, const double* __restrict__ b,
                                                  double* __restrict__ q, double* part, int max_blocks) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    double v[2] = {0.0, 0.0};
    if (i < D.ni && j < D.nj) {
        const size_t p = D.off(i, j);
        const size_t ld = D.ld;
        const double wp = w[p];
        const double qv =
            diag[p] * wp - cE[p - 1] * w[p - 1] - cE[p] * w[p + 1] - cN[p - ld] * w[p - ld] - cN[p] * w[p + ld];
        q[p] = qv;
        v[0] = wp * b[p];
        v[1] = wp * qv;
    }
    block_partials<2, 256>(v, part, max_blocks);
}

__global__ void __launch_bounds__(NTR) k_galerkin_fin(const double* part, int max_blocks, int nblocks, PcgScalars* sc) {
    double s[2];
    finalize_sum<2, NTR>(part, max_blocks, nblocks, s);
    if (threadIdx.x != 0) return;
    sc->wb = s[0];
    sc->wq = s[1];
    sc->walpha = s[1] > 0.0 ? s[0] / s[1] : 0.0;
}

// x = walpha * w, or x += walpha * w on top of the previous solution
__global__ void __launch_bounds__(256) k_start(Dims D, const double* __restrict__ w, const PcgScalars* sc,
                                               double* __restrict__ x, int add) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= D.ni || j >= D.nj) return;
    const size_t p = D.off(i, j);
    x[p] = add ? x[p] + sc->walpha * w[p] : sc->walpha * w[p];
}

// r = b - A x on the fine level (the residual of the previous solution)
__global__ void __launch_bounds__(256) k_resid_prev(Dims D, const double* __restrict__ diag,
                                                    const double* __restrict__ cE, const double* __restrict__ cN,
                                                    const double* __restrict__ b, const double* __restrict__ x,
                                                    double* __restrict__ r) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= D.ni || j >= D.nj) return;
    const size_t p = D.off(i, j);
    const size_t ld = D.ld;
    const double ax =
        diag[p] * x[p] - cE[p - 1] * x[p - 1] - cE[p] * x[p + 1] - cN[p - ld] * x[p - ld] - cN[p] * x[p + ld];
    r[p] = b[p] - ax;
}

bool net_fits(const Dims& d) { return d.nj % 8 == 0 && d.ni % 128 == 0; }

int restrict_to(Ctx* c, int level, const double* src) {
    hipStream_t s = c->stream;
    for (int l = 0; l < level; ++l) {
        const Dims& C = c->lev[l + 1].d;
        hipLaunchKernelGGL(k_restrict, grid2d(C.ni, C.nj), block2d(), 0, s, c->lev[l].d, C, c->lev[l].P, src,
                           c->lev[l + 1].v.b, (l == 0 && c->semi0) ? 1 : 0);
        src = c->lev[l + 1].v.b;
    }
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int features(Ctx* c, int level, bf16* X, float* t, float* mask, const double* src) {
    int rc = restrict_to(c, level, src);
    if (rc) return rc;
    Level& L = c->lev[level];
    const double* b = level ? L.v.b : src;
    {
        const dim3 g = grid2d(L.d.ni, L.d.nj);
        double* part = c->red.partials + (size_t)5 * NRED * c->red.max_blocks;
        hipLaunchKernelGGL(k_rms, g, block2d(), 0, c->stream, L.d, (const double*)L.A.diag, b, part,
                           c->red.max_blocks);
        hipLaunchKernelGGL(k_rms_fin, dim3(1), dim3(NTR), 0, c->stream, (const double*)part, c->red.max_blocks,
                           (int)(g.x * g.y), c->sc);
    }
    hipLaunchKernelGGL(k_features, grid2d(L.d.ni, L.d.nj), block2d(), 0, c->stream, L.d, L.A, b, c->lev[0].d,
                       (const double*)c->field[SG_F_VOF], c->ni, c->nj, level, c->semi0, (const double*)c->x,
                       (const PcgScalars*)c->sc, X, t, mask);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

}  // namespace

void learned_release(Ctx* c) {
    for (hipGraphExec_t* ge : {&c->nn.graph, &c->nn.graph_prev})
        if (*ge) {
            (void)hipGraphExecDestroy(*ge);
            *ge = nullptr;
        }
    if (c->nn.net) {
        nn::destroy(c->nn.net);
        delete c->nn.net;
        c->nn.net = nullptr;
    }
    c->nn.on = 0;
}

namespace {
// the fine-level buffers both the net and a caller's prediction carry to
int learned_buffers(Ctx* c) {
    if (!c->nn.X) {
        // input buffers sized for the largest admissible level (the fine grid)
        const size_t xb = (size_t)(c->nj + 2) * (c->ni + 2) * 8 * sizeof(bf16);
        const size_t fb = c->lev[0].d.alloc_elems() * sizeof(double);
        c->nn.X = dev_alloc(c, xb);
        c->nn.w = (double*)dev_alloc(c, fb);
        c->nn.aw = (double*)dev_alloc(c, fb);
        if (!c->nn.X || !c->nn.w || !c->nn.aw) return SG_ERR_ALLOC;
        SG_HIP(hipMemset(c->nn.X, 0, xb));
        SG_HIP(hipMemset(c->nn.w, 0, fb));
        SG_HIP(hipMemset(c->nn.aw, 0, fb));
    }
    if (!c->ev_nn0) {
        SG_HIP(hipEventCreate(&c->ev_nn0));
        SG_HIP(hipEventCreate(&c->ev_nn1));
    }
    return SG_OK;
}
}  // namespace

int learned_set_prediction(Ctx* c, int level, const double* e_host, int deflate) {
    if (!e_host) {
        c->nn.ext_on = 0;
        return SG_OK;
    }
    if (level < 0 || level >= c->nlev) {
        set_error("sg_set_learned_prediction: no such level");
        return SG_ERR_ARG;
    }
    int rc = learned_buffers(c);
    if (rc) return rc;
    if (!c->nn.ext) {
        // sized for the fine level, so any level's prediction fits
        const size_t fb = c->lev[0].d.alloc_elems() * sizeof(double);
        c->nn.ext = (double*)dev_alloc(c, fb);
        if (!c->nn.ext) return SG_ERR_ALLOC;
    }
    const Dims& d = c->lev[level].d;
    SG_HIP(hipMemsetAsync(c->nn.ext, 0, d.alloc_elems() * sizeof(double), c->stream));
    SG_HIP(hipMemcpy2DAsync(c->nn.ext + d.off(0, 0), d.ld * sizeof(double), e_host, d.ni * sizeof(double),
                            d.ni * sizeof(double), d.nj, hipMemcpyHostToDevice, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    learned_release(c);  // the prediction replaces a net
    c->nn.level = level;
    c->nn.out_scale = 1.0;
    c->nn.deflate = deflate ? 1 : 0;
    c->nn.ext_on = 1;
    return SG_OK;
}

int learned_enable(Ctx* c, const float* w, int level, double out_scale, int deflate) {
    learned_release(c);
    if (!w) return SG_OK;
    c->nn.ext_on = 0;
    if (level < 0 || level >= c->nlev || !net_fits(c->lev[level].d)) {
        set_error("sg_set_learned: the level's grid must exist and have nj % 8 == 0, ni % 128 == 0");
        return SG_ERR_ARG;
    }
    const Dims& d = c->lev[level].d;
    c->nn.net = new nn::Net();
    int rc = nn::create(c->nn.net, d.nj, d.ni, 1, c->stream);
    if (rc) {
        delete c->nn.net;
        c->nn.net = nullptr;
        return rc;
    }
    rc = nn::set_weights(c->nn.net, w);
    if (rc) {
        learned_release(c);
        return rc;
    }
    rc = learned_buffers(c);
    if (rc) {
        learned_release(c);
        return rc;
    }
    c->nn.level = level;
    c->nn.out_scale = out_scale;
    c->nn.deflate = deflate ? 1 : 0;
    c->nn.on = 1;
    return SG_OK;
}

int learned_features(Ctx* c, int level, void* X, float* t, float* mask, double* scale) {
    if (!c->mg_ready) {
        set_error("sg_nn_features: needs a solve of the current system first");
        return SG_ERR_ARG;
    }
    if (level < 0 || level >= c->nlev) {
        set_error("sg_nn_features: no such level");
        return SG_ERR_ARG;
    }
    int rc = features(c, level, (bf16*)X, t, mask, c->rhs);
    if (rc) return rc;
    SG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    if (scale) *scale = c->sc_host->nn_scale;
    return SG_OK;
}

namespace {

// x0 = (w,b)/(w,Aw) w with w the net's prediction carried to the fine grid
int enqueue_warm_start(Ctx* c, bool prev) {
    hipStream_t s = c->stream;
    const int L = c->nn.level;
    // from the previous solution: the net predicts the correction from the
    // residual that solution leaves
    const double* src = c->rhs;
    if (prev) {
        const Dims D0 = c->lev[0].d;
        hipLaunchKernelGGL(k_resid_prev, grid2d(c->ni, c->nj), block2d(), 0, s, D0,
                           (const double*)c->lev[0].A.diag, (const double*)c->lev[0].A.cE,
                           (const double*)c->lev[0].A.cN, (const double*)c->rhs, (const double*)c->x, c->r);
        src = c->r;
    }
    Level& Lv = c->lev[L];
    double* e = L ? Lv.v.x : c->nn.w;
    if (c->nn.ext_on) {
        // the caller's level-L prediction in place of the net's
        hipLaunchKernelGGL(k_take_prediction, grid2d(Lv.d.ni, Lv.d.nj), block2d(), 0, s, Lv.d,
                           (const double*)Lv.A.diag, (const double*)c->nn.ext, e);
    } else {
        int rc = features(c, L, (bf16*)c->nn.X, nullptr, nullptr, src);
        if (rc) return rc;
        rc = nn::forward(c->nn.net, (const bf16*)c->nn.X, 1);
        if (rc) return rc;
        hipLaunchKernelGGL(k_from_net, grid2d(Lv.d.ni, Lv.d.nj), block2d(), 0, s, Lv.d, (const double*)Lv.A.diag,
                           (const float*)c->nn.net->y, c->nn.out_scale, (const PcgScalars*)c->sc, e);
    }
    for (int l = L - 1; l >= 0; --l) {
        double* dst = l ? c->lev[l].v.x : c->nn.w;
        hipLaunchKernelGGL(k_prolong, grid2d(c->lev[l].d.ni, c->lev[l].d.nj), block2d(), 0, s, c->lev[l].d,
                           c->lev[l + 1].d, c->lev[l].P, (const double*)c->lev[l].A.diag,
                           (const double*)c->lev[l + 1].v.x, dst, (l == 0 && c->semi0) ? 1 : 0);
    }
    const Dims D = c->lev[0].d;
    const dim3 g = grid2d(c->ni, c->nj);
    const int nb = (int)(g.x * g.y);
    double* part = c->red.partials + (size_t)5 * NRED * c->red.max_blocks;
    hipLaunchKernelGGL(k_galerkin, g, block2d(), 0, s, D, (const double*)c->lev[0].A.diag,
                       (const double*)c->lev[0].A.cE, (const double*)c->lev[0].A.cN, (const double*)c->nn.w,
                       src, c->nn.aw, part, c->red.max_blocks);
    hipLaunchKernelGGL(k_galerkin_fin, dim3(1), dim3(NTR), 0, s, (const double*)part, c->red.max_blocks, nb, c->sc);
    hipLaunchKernelGGL(k_start, g, block2d(), 0, s, D, (const double*)c->nn.w, (const PcgScalars*)c->sc, c->x,
                       prev ? 1 : 0);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

}  // namespace

// The warm start's ~45 kernels (restriction, features, the net, the
// prolongation, the Galerkin scaling) run as one graph, captured at the first
// call after sg_set_learned: every buffer it touches lives as long as the net.
int learned_warm_start(Ctx* c, bool prev) {
    hipStream_t s = c->stream;
    if (c->nn.ext_on) {
        SG_HIP(hipEventRecord(c->ev_nn0, s));
        const int rc = enqueue_warm_start(c, prev);
        if (rc) return rc;
        SG_HIP(hipEventRecord(c->ev_nn1, s));
        return SG_OK;
    }
    hipGraphExec_t& ge = prev ? c->nn.graph_prev : c->nn.graph;
    if (!ge) {
        hipGraph_t graph = nullptr;
        SG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_warm_start(c, prev);
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        const hipError_t e2 = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e2 != hipSuccess) {
            ge = nullptr;
            return hip_fail(e2, "hipGraphInstantiate");
        }
    }
    SG_HIP(hipEventRecord(c->ev_nn0, s));
    SG_HIP(hipGraphLaunch(ge, s));
    SG_HIP(hipEventRecord(c->ev_nn1, s));
    return SG_OK;
}

}  // namespace sg
