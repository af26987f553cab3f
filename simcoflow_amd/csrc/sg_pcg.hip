// Preconditioned conjugate gradients for the projection (stands where
// HYPRE_ParCSRPCGSolve stands, ProjectionP_Mod.f90:57-58) with a multigrid
// V-cycle preconditioner.  All scalars (alpha, beta, norms) stay on the
// device; the host reads them only to test convergence.
//
// FP64: every vector and the operator in FP64.  Mixed precision: the PCG
// vectors r, z, p, q and the whole cycle are FP32, q = A p applies the FP64
// operator; x and b stay FP64, every dot product is accumulated in FP64, and the
// residual is recomputed from x in FP64 (r = b - A x, "reliable update")
// whenever the recursive one has fallen by 100x since the last
// recomputation and again at convergence, so the FP32 recursion never
// limits the attained FP64 residual.  The FP32 cycle is not exactly
// symmetric, so beta takes the Polak-Ribiere (flexible) form
// -alpha*(z_new,q_old)/(r_old,z_old).  The recomputation test runs on the
// device, inside the captured iterations, so the host only tests
// convergence -- after a number of iterations it sets from the observed
// rate; iterations launched past convergence cost one flag test per kernel.
//
// Deflation (a learned vector w): every new direction p = z + beta p_old -
// mu w is re-projected A-orthogonal to w after q = A p is formed, p' = p -
// gamma w, q' = q - gamma A w with gamma = (w, A p) / (w, A w); the
// correction is folded into the update (x += alpha p', r -= alpha q') and
// into the next direction (its w coefficient mu + beta gamma), so an FP32
// direction cannot drift away from A-orthogonality over the iterations.
// gamma takes A p as the FP64 stencil sums, before q is rounded to the
// vector precision: then (A w, p') vanishes to FP64 rounding, x gains no
// component along w, and the FP64 residual recomputed from x stays
// orthogonal to w.  (With the rounded q, (w, A p') is of the order of the
// FP32 rounding of q at every step; it accumulates in x, no A-orthogonal
// direction can remove it, and the recomputed residual then carries a w
// part the FP32 recursion stalls and diverges on.)
#include "sg_internal.hpp"
#include "sg_reduce.hpp"

namespace sg {

namespace {

constexpr int NT = TX * TY;
constexpr int NTF = 1024;  // threads of the finalize kernels

enum FinMode {
    FIN_INIT = 0, FIN_REPLACE, FIN_RZ, FIN_PQ, FIN_RR, FIN_TRUE, FIN_MU, FIN_RELIABLE, FIN_RELIABLE_IF, FIN_PQ_DEFL
};

// mixed precision: the FP32 residual is recomputed from x in FP64 once it
// has fallen by this factor (in (r,r)) since the last recomputation
constexpr double RELIABLE_DROP = 1e-4;

__device__ inline bool reliable_due(const PcgScalars* sc) {
    return sc->active && sc->rr < RELIABLE_DROP * sc->rr_ref;
}

// one reduction slot: NRED planes of per-block partials
struct Slots {
    double* part;
    int max_blocks;
};

// r = b - A x in FP64, stored as V (or r = b for a zero start); partials of
// ||b||^2, ||r||^2.  With `gate`, only when a recomputation is due.
template <typename V>
__global__ void __launch_bounds__(256) k_init(Dims D, const double* __restrict__ diag, const double* __restrict__ cE,
                                              const double* __restrict__ cN, const double* __restrict__ b,
                                              double* __restrict__ x, V* __restrict__ r, int warm, Slots S,
                                              const PcgScalars* gate) {
    if (gate && !reliable_due(gate)) return;
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    double v[2] = {0.0, 0.0};
    if (i < D.ni && j < D.nj) {
        const size_t p = D.off(i, j);
        const size_t ld = D.ld;
        const double bp = b[p];
        double rp = bp;
        if (warm) {
            // a previous solution is zero on solid rows already
            rp = bp - (diag[p] * x[p] - cE[p - 1] * x[p - 1] - cE[p] * x[p + 1] - cN[p - ld] * x[p - ld] -
                       cN[p] * x[p + ld]);
            if (!(diag[p] > 0.0)) rp = 0.0;
        } else {
            x[p] = 0.0;
        }
        r[p] = (V)rp;
        v[0] = bp * bp;
        v[1] = rp * rp;
    }
    block_partials<2, NT>(v, S.part, S.max_blocks);
}

// partials of rz = (r,z), zq = (z,q_old)
template <typename V>
__global__ void __launch_bounds__(256) k_dots_rz(Dims D, const V* __restrict__ r, const V* __restrict__ z,
                                                 const V* __restrict__ q, Slots S, const PcgScalars* sc) {
    if (!sc->active) return;
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    double v[2] = {0.0, 0.0};
    if (i < D.ni && j < D.nj) {
        const size_t p = D.off(i, j);
        const double zp = (double)z[p];
        v[0] = (double)r[p] * zp;
        v[1] = zp * (double)q[p];
    }
    block_partials<2, NT>(v, S.part, S.max_blocks);
}

// partials of (A w, z): the deflation coefficient mu = (A w, z) / (w, A w)
template <typename V>
__global__ void __launch_bounds__(256) k_dot_mu(Dims D, const double* __restrict__ aw, const V* __restrict__ z,
                                                Slots S, const PcgScalars* sc) {
    if (!sc->active) return;
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    double v[1] = {0.0};
    if (i < D.ni && j < D.nj) {
        const size_t p = D.off(i, j);
        v[0] = aw[p] * (double)z[p];
    }
    block_partials<1, NT>(v, S.part, S.max_blocks);
}

// p_new = z + beta p_old (- mu w with deflation) ; q = A p_new ;
// partials of (p_new, q).  A 64x16 tile: p_new of the tile and its one-cell
// halo is formed once in LDS, then the stencil reads it from there.
// V = type of the vectors (FP64, or FP32 in mixed precision); the operator
// is the FP64 one in both (an FP32 copy would make the recursion solve a
// perturbed system, which the FP64 residual recomputation then has to undo
// at a rate set by the condition number), the stencil sums and (p,q) in
// FP64.
constexpr int DW = 64, DH = 16;
template <typename V>
__global__ void __launch_bounds__(256) k_dir_apply(Dims D, const double* __restrict__ diag,
                                                   const double* __restrict__ cE, const double* __restrict__ cN,
                                                   const V* __restrict__ z, const V* __restrict__ pold,
                                                   const double* __restrict__ w, V* __restrict__ pnew,
                                                   V* __restrict__ q, Slots S, const PcgScalars* sc) {
    if (!sc->active) return;
    constexpr int RW = DW + 2, RH = DH + 2;
    __shared__ V sp[RH][RW];
    const V beta = (V)sc->beta;
    const V mu = w ? (V)sc->mu_eff : V(0);
    // the first direction reads no old one: p_old still holds the last
    // solve's, which may not be finite after a failed solve (0 * NaN)
    const bool first = sc->first != 0;
    int tx, ty;
    xcd_tile(tx, ty);
    const int i0 = tx * DW, j0 = ty * DH;
    const int tid = threadIdx.x;
    for (int k = tid; k < RW * RH; k += 256) {
        const int li = k % RW, lj = k / RW;
        const int gi = i0 + li - 1, gj = j0 + lj - 1;
        V v = V(0);
        if (gi >= 0 && gj >= 0 && gi < D.ni && gj < D.nj) {
            const size_t p = D.off(gi, gj);
            const V zb = first ? z[p] : z[p] + beta * pold[p];
            v = w ? zb - mu * (V)w[p] : zb;
        }
        sp[lj][li] = v;
    }
    __syncthreads();
    double v[2] = {0.0, 0.0};
    const size_t ld = D.ld;
#pragma unroll
    for (int m = 0; m < DW * DH / 256; ++m) {
        const int k = tid + m * 256;
        const int li = k % DW + 1, lj = k / DW + 1;
        const int gi = i0 + li - 1, gj = j0 + lj - 1;
        if (gi >= D.ni || gj >= D.nj) continue;
        const size_t p = D.off(gi, gj);
        const V pc = sp[lj][li];
        const double qd = diag[p] * (double)pc - cE[p - 1] * (double)sp[lj][li - 1] -
                          cE[p] * (double)sp[lj][li + 1] - cN[p - ld] * (double)sp[lj - 1][li] -
                          cN[p] * (double)sp[lj + 1][li];
        const V qv = (V)qd;
        pnew[p] = pc;
        q[p] = qv;
        v[0] += (double)pc * (double)qv;
        if (w) v[1] += w[p] * qd;  // (w, A p) before rounding
    }
    block_partials<2, 256>(v, S.part, S.max_blocks);
}

// x += alpha p (x in FP64) ; r -= alpha q ; partials of (r,r)
// with deflation (w, aw = A w): the re-projected p - gamma w, q - gamma A w
template <typename V>
__global__ void __launch_bounds__(256) k_update(Dims D, const V* __restrict__ p, const V* __restrict__ q,
                                                const double* __restrict__ w, const double* __restrict__ aw,
                                                double* __restrict__ x, V* __restrict__ r, Slots S,
                                                const PcgScalars* sc) {
    if (!sc->active) return;
    const double alpha = sc->alpha;
    const V alv = (V)alpha;
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    double v[1] = {0.0};
    if (i < D.ni && j < D.nj) {
        const size_t k = D.off(i, j);
        if (w) {
            const double gm = sc->gamma;
            x[k] += alpha * ((double)p[k] - gm * w[k]);
            const V rn = (V)((double)r[k] - alpha * ((double)q[k] - gm * aw[k]));
            r[k] = rn;
            v[0] = (double)rn * (double)rn;
        } else {
            x[k] += alpha * (double)p[k];
            const V rn = r[k] - alv * q[k];
            r[k] = rn;
            v[0] = (double)rn * (double)rn;
        }
    }
    block_partials<1, NT>(v, S.part, S.max_blocks);
}

// partials of the true residual: sum of squares and max |.| over fluid rows
__global__ void __launch_bounds__(256) k_true_resid(Dims D, const double* __restrict__ diag,
                                                    const double* __restrict__ cE, const double* __restrict__ cN,
                                                    const double* __restrict__ b, const double* __restrict__ x,
                                                    Slots S) {
    __shared__ double smax[NT / 64];
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    double v[1] = {0.0};
    double m = 0.0;
    if (i < D.ni && j < D.nj) {
        const size_t p = D.off(i, j);
        const size_t ld = D.ld;
        if (diag[p] > 0.0) {
            const double rr = b[p] - (diag[p] * x[p] - cE[p - 1] * x[p - 1] - cE[p] * x[p + 1] -
                                      cN[p - ld] * x[p - ld] - cN[p] * x[p + ld]);
            v[0] = rr * rr;
            m = fabs(rr);
        }
    }
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if ((tid & 63) == 0) smax[tid >> 6] = m;
    block_partials<1, NT>(v, S.part, S.max_blocks);  // synchronises the block
    if (tid == 0) {
        double c = 0.0;
        for (int w = 0; w < NT / 64; ++w) c = fmax(c, smax[w]);
        S.part[(size_t)S.max_blocks + blockIdx.x + blockIdx.y * gridDim.x] = c;
    }
}

// One block: reduce the partials in block order and update the scalars.
__global__ void __launch_bounds__(NTF) k_finalize(Slots S, int nblocks, int mode, PcgScalars* sc, double tol) {
    if (mode == FIN_RELIABLE_IF) {
        if (!reliable_due(sc)) return;
        mode = FIN_RELIABLE;
    }
    if (mode != FIN_INIT && mode != FIN_REPLACE && mode != FIN_TRUE && mode != FIN_RELIABLE && !sc->active) return;
    if (mode == FIN_TRUE) {
        __shared__ double mx[NTF / 64];
        double s[1];
        finalize_sum<1, NTF>(S.part, S.max_blocks, nblocks, s);
        double m = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += NTF) m = fmax(m, S.part[(size_t)S.max_blocks + b]);
        for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
        if ((threadIdx.x & 63) == 0) mx[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            double c = 0.0;
            for (int w = 0; w < NTF / 64; ++w) c = fmax(c, mx[w]);
            sc->true_rr = s[0];
            sc->true_max = c;
        }
        return;
    }
    if (mode == FIN_RELIABLE) {
        double s2[2];
        finalize_sum<2, NTF>(S.part, S.max_blocks, nblocks, s2);
        if (threadIdx.x != 0) return;
        sc->rr = s2[1];
        sc->rr_ref = s2[1];
        sc->active = (s2[1] > sc->tol2) ? 1 : 0;
        if (isnan(s2[1])) {
            sc->active = 0;
            sc->guard |= 1;
        }
        return;
    }
    if (mode == FIN_INIT || mode == FIN_REPLACE || mode == FIN_RZ) {
        double s[2];
        finalize_sum<2, NTF>(S.part, S.max_blocks, nblocks, s);
        if (threadIdx.x != 0) return;
        if (mode == FIN_RZ) {
            double beta = 0.0;
            if (!sc->first) beta = sc->flexible ? -sc->alpha * s[1] / sc->rz_old : s[0] / sc->rz_old;
            sc->rz = s[0];
            sc->zq = s[1];
            sc->beta = beta;
            return;
        }
        // FIN_INIT / FIN_REPLACE: s[0] = ||b||^2, s[1] = ||r||^2
        if (mode == FIN_INIT) {
            sc->bb = s[0];
            sc->tol2 = tol * tol * s[0];
            sc->iters = 0;
        }
        sc->rr = s[1];
        sc->rr_ref = s[1];
        sc->first = 1;
        sc->beta = 0.0;
        sc->gamma = 0.0;
        sc->active = (s[0] > 0.0 && s[1] > sc->tol2) ? 1 : 0;
        sc->rz = sc->rz_old = sc->zq = sc->pq = 0.0;
        return;
    }
    if (mode == FIN_PQ || mode == FIN_PQ_DEFL) {
        double s2[2];
        finalize_sum<2, NTF>(S.part, S.max_blocks, nblocks, s2);
        if (threadIdx.x != 0) return;
        double pq = s2[0];
        if (mode == FIN_PQ_DEFL) {
            // re-projection of the direction A-orthogonal to w: (p', q') =
            // (p, q) - gamma (w, A p)
            const double gm = sc->wq > 0.0 ? s2[1] / sc->wq : 0.0;
            sc->gamma = gm;
            pq = s2[0] - gm * s2[1];
        }
        sc->pq = pq;
        sc->alpha = sc->rz / pq;
        sc->first = 0;
        return;
    }
    double s[1];
    finalize_sum<1, NTF>(S.part, S.max_blocks, nblocks, s);
    if (threadIdx.x != 0) return;
    if (mode == FIN_MU) {
        // s[0] = (A w, z).  The last direction was used as p' = p - gamma w,
        // q' = q - gamma A w: the flexible beta reads (z, q') = (z, q) -
        // gamma (A w, z), and the new direction's w coefficient takes
        // beta gamma beside mu.
        sc->mu = sc->wq > 0.0 ? s[0] / sc->wq : 0.0;
        if (!sc->first && sc->flexible) sc->beta = -sc->alpha * (sc->zq - sc->gamma * s[0]) / sc->rz_old;
        sc->mu_eff = sc->mu + (sc->first ? 0.0 : sc->beta * sc->gamma);
    } else {  // FIN_RR
        sc->rr = s[0];
        sc->rz_old = sc->rz;
        sc->iters += 1;
        if (!(s[0] > sc->tol2)) sc->active = 0;
        if (isnan(s[0])) {
            sc->active = 0;
            sc->guard |= 1;
        }
    }
}

Slots slots(Ctx* c, int s) {
    return Slots{c->red.partials + (size_t)s * NRED * c->red.max_blocks, c->red.max_blocks};
}

bool learned(Ctx* c) { return (c->prm.warm_start == 2 || c->prm.warm_start == 3) && (c->nn.on || c->nn.ext_on); }
bool deflating(Ctx* c) { return learned(c) && c->nn.deflate; }

// one PCG iteration (V-cycle + direction + update), enqueued on the stream;
// V = the vector precision, A the operator of that precision
template <typename V>
int enqueue_iteration_t(Ctx* c, int parity, V* r, V* z, V* q, V* p0, V* p1) {
    const Op<double>& A = c->lev[0].A;
    const Dims D = c->lev[0].d;
    hipStream_t s = c->stream;
    const dim3 g = grid2d(c->ni, c->nj), b = block2d();
    const int nb = (int)(g.x * g.y);
    V* pold = parity ? p1 : p0;
    V* pnew = parity ? p0 : p1;
    Slots S1 = slots(c, 1), S2 = slots(c, 2), S3 = slots(c, 3);
    // the cycle's last kernel also leaves the partials of (r,z) and (z,q)
    int dots_done = 0;
    int rc = mg_vcycle_dots(c, S1.part, S1.max_blocks, &dots_done);
    if (rc) return rc;
    int nb_rz = mg_partial_blocks(c);
    if (!dots_done) {
        hipLaunchKernelGGL(k_dots_rz<V>, g, b, 0, s, D, (const V*)r, (const V*)z, (const V*)q, S1,
                           (const PcgScalars*)c->sc);
        nb_rz = nb;
    }
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(NTF), 0, s, S1, nb_rz, (int)FIN_RZ, c->sc, 0.0);
    const double* w = nullptr;
    if (deflating(c)) {
        Slots S6 = slots(c, 6);
        hipLaunchKernelGGL(k_dot_mu<V>, g, b, 0, s, D, (const double*)c->nn.aw, (const V*)z, S6,
                           (const PcgScalars*)c->sc);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(NTF), 0, s, S6, nb, (int)FIN_MU, c->sc, 0.0);
        w = c->nn.w;
    }
    const dim3 gd((c->ni + DW - 1) / DW, (c->nj + DH - 1) / DH);
    hipLaunchKernelGGL(k_dir_apply<V>, gd, dim3(256), 0, s, D, (const double*)A.diag, (const double*)A.cE,
                       (const double*)A.cN, (const V*)z, (const V*)pold, w, pnew, q, S2, (const PcgScalars*)c->sc);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(NTF), 0, s, S2, (int)(gd.x * gd.y), (int)(w ? FIN_PQ_DEFL : FIN_PQ),
                       c->sc, 0.0);
    hipLaunchKernelGGL(k_update<V>, g, b, 0, s, D, (const V*)pnew, (const V*)q, w, w ? (const double*)c->nn.aw : nullptr,
                       c->x, r, S3, (const PcgScalars*)c->sc);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(NTF), 0, s, S3, nb, (int)FIN_RR, c->sc, 0.0);
    if (sizeof(V) == 4 && parity == 1) {
        // FP32 vectors: recompute r from x in FP64 when it is due, keeping
        // the search direction
        Slots S0 = slots(c, 0);
        hipLaunchKernelGGL(k_init<V>, g, b, 0, s, D, (const double*)A.diag, (const double*)A.cE,
                           (const double*)A.cN, (const double*)c->rhs, c->x, r, 1, S0, (const PcgScalars*)c->sc);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(NTF), 0, s, S0, nb, (int)FIN_RELIABLE_IF, c->sc, 0.0);
    }
    SG_HIP(hipGetLastError());
    return SG_OK;
}

bool mixed(Ctx* c) { return c->prm.precision == SG_PREC_MIXED; }

int enqueue_iteration(Ctx* c, int parity) {
    if (mixed(c)) return enqueue_iteration_t<float>(c, parity, c->rf, c->zf, c->qf, c->pf0, c->pf1);
    return enqueue_iteration_t<double>(c, parity, c->r, c->z, c->q, c->p0, c->p1);
}

// r = b - A x (warm) or r = b, x = 0 (cold), in FP64, stored at the
// context's vector precision; then the finalize of `mode`
int enqueue_init(Ctx* c, int warm, int mode) {
    Level& L0 = c->lev[0];
    const dim3 g = grid2d(c->ni, c->nj);
    const int nb = (int)(g.x * g.y);
    Slots S0 = slots(c, 0);
    if (mixed(c))
        hipLaunchKernelGGL(k_init<float>, g, block2d(), 0, c->stream, L0.d, L0.A.diag, L0.A.cE, L0.A.cN,
                           (const double*)c->rhs, c->x, c->rf, warm, S0, nullptr);
    else
        hipLaunchKernelGGL(k_init<double>, g, block2d(), 0, c->stream, L0.d, L0.A.diag, L0.A.cE, L0.A.cN,
                           (const double*)c->rhs, c->x, c->r, warm, S0, nullptr);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(NTF), 0, c->stream, S0, nb, mode, c->sc, c->prm.tol);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

// PCG iterations captured once into graphs: variant 0 = GRAPH_ITERS
// iterations (parities 0, 1, 0, 1, ...), 1 = one iteration of parity 0, 2 =
// one of parity 1.  Each graph launch costs a pause of tens of microseconds
// partway through its first iteration, so the long graph carries several
// iterations.
constexpr int GRAPH_ITERS = 4;
static_assert(GRAPH_ITERS % 2 == 0, "the long graph must end on the parity it starts from");
int iteration_graph(Ctx* c, int variant, hipGraphExec_t* out) {
    const int k = c->prm.precision == SG_PREC_MIXED ? 1 : 0;
    const int defl = deflating(c) ? 1 : 0;
    if (c->graph_pre[k] != c->prm.mg_pre || c->graph_post[k] != c->prm.mg_post || c->graph_defl[k] != defl ||
        c->graph_kry[k] != c->prm.mg_coarse_krylov) {
        for (hipGraphExec_t& g : c->iter_graph[k]) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
        c->graph_pre[k] = c->prm.mg_pre;
        c->graph_post[k] = c->prm.mg_post;
        c->graph_defl[k] = defl;
        c->graph_kry[k] = c->prm.mg_coarse_krylov;
    }
    if (c->iter_graph[k][variant]) {
        *out = c->iter_graph[k][variant];
        return SG_OK;
    }
    hipGraph_t graph = nullptr;
    SG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_iteration(c, variant == 2 ? 1 : 0);
    for (int q = 1; rc == SG_OK && variant == 0 && q < GRAPH_ITERS; ++q) rc = enqueue_iteration(c, q & 1);
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    c->iter_graph[k][variant] = exec;
    *out = exec;
    return SG_OK;
}

// enqueue n PCG iterations, continuing the direction parity c->pcur
int launch_iterations(Ctx* c, int n) {
    while (n > 0) {
        if (c->use_graph) {
            const int variant = (c->pcur == 0 && n >= GRAPH_ITERS) ? 0 : (c->pcur == 0 ? 1 : 2);
            hipGraphExec_t g = nullptr;
            const int rc = iteration_graph(c, variant, &g);
            if (rc) return rc;
            SG_HIP(hipGraphLaunch(g, c->stream));
            if (variant == 0) {
                n -= GRAPH_ITERS;
            } else {
                n -= 1;
                c->pcur ^= 1;
            }
        } else {
            const int rc = enqueue_iteration(c, c->pcur);
            if (rc) return rc;
            c->pcur ^= 1;
            n -= 1;
        }
    }
    return SG_OK;
}

}  // namespace

int fine_residual_norms(Ctx* c, double* relres, double* maxrow) {
    Level& L0 = c->lev[0];
    const dim3 g = grid2d(c->ni, c->nj);
    const int nb = (int)(g.x * g.y);
    Slots S = slots(c, 4);
    hipLaunchKernelGGL(k_true_resid, g, block2d(), 0, c->stream, L0.d, L0.A.diag, L0.A.cE, L0.A.cN,
                       (const double*)c->rhs, (const double*)c->x, S);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(NTF), 0, c->stream, S, nb, (int)FIN_TRUE, c->sc, 0.0);
    SG_HIP(hipGetLastError());
    SG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    const double bb = c->sc_host->bb;
    *relres = bb > 0.0 ? sqrt(c->sc_host->true_rr / bb) : 0.0;
    *maxrow = c->sc_host->true_max;
    return SG_OK;
}

int pcg_solve(Ctx* c, sg_solve_info* info) {
    hipStream_t s = c->stream;
    SG_HIP(hipEventRecord(c->ev0, s));
    if (!c->mg_ready) {
        int rc = mg_setup(c);
        if (rc) return rc;
    }
    SG_HIP(hipEventRecord(c->ev1, s));
    // flexible beta in mixed precision and with inner iterations on the
    // coarse levels (both make the preconditioner vary slightly)
    c->sc_host->flexible = (c->prm.precision == SG_PREC_MIXED || c->prm.mg_coarse_krylov > 0) ? 1 : 0;
    SG_HIP(hipMemcpyAsync(&c->sc->flexible, &c->sc_host->flexible, sizeof(int), hipMemcpyHostToDevice, s));
    int rc;
    if (c->prm.warm_start == 2 || c->prm.warm_start == 3) {
        if (!c->nn.on && !c->nn.ext_on) {
            set_error("warm_start = 2 or 3 needs the learned warm start (sg_set_learned or sg_set_learned_prediction)");
            return SG_ERR_ARG;
        }
        rc = learned_warm_start(c, c->prm.warm_start == 3);
        if (rc) return rc;
        rc = enqueue_init(c, 1, FIN_INIT);
    } else {
        rc = enqueue_init(c, c->prm.warm_start, FIN_INIT);
    }
    if (rc) return rc;
    SG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    const double warm_relres = c->sc_host->bb > 0.0 ? sqrt(c->sc_host->rr / c->sc_host->bb) : 0.0;
    float warm_ms = 0.f;
    double warm_energy = 0.0;
    if (c->prm.warm_start == 2 || c->prm.warm_start == 3) {
        SG_HIP(hipEventElapsedTime(&warm_ms, c->ev_nn0, c->ev_nn1));
        if (c->sc_host->wq > 0.0) warm_energy = c->sc_host->wb * c->sc_host->wb / c->sc_host->wq;
    }
    const int max_iter = c->prm.max_iter > 0 ? c->prm.max_iter : 1000;
    const int every = c->prm.check_every;
    c->pcur = 0;
    int it = 0;  // iterations launched (those past convergence do nothing)
    int replacements = 0;
    // convergence rate per iteration (log of the (r,r) ratio) between host
    // checks; the last solve's to begin with (the systems of consecutive
    // time steps converge alike, so the first launch usually reaches the
    // tolerance), 0 while unknown.  Iterations past convergence do nothing,
    // so the launch sizes never change the result.
    double rate = c->pcg_rate;
    for (;;) {
        while (c->sc_host->active && it < max_iter) {
            int n = 4;  // before the rate is known
            if (every > 0) {
                n = every;
            } else if (rate < 0.0) {
                const double need = log(c->sc_host->tol2 / c->sc_host->rr) / rate;
                n = need < 1.0 ? 1 : (need > 16.0 ? 16 : (int)ceil(need));
            }
            if (n > max_iter - it) n = max_iter - it;
            const double rr0 = c->sc_host->rr;
            const int k0 = c->sc_host->iters;
            rc = launch_iterations(c, n);
            if (rc) return rc;
            it += n;
            SG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
            SG_HIP(hipStreamSynchronize(s));
            const int done = c->sc_host->iters - k0;
            if (done > 0 && c->sc_host->rr > 0.0 && c->sc_host->rr < rr0) {
                rate = log(c->sc_host->rr / rr0) / done;
                c->pcg_rate = rate;
            }
        }
        // residual replacement at convergence: the recursive residual drifts
        // from b - A x by rounding; recompute it in FP64 from x and restart CG
        // from it until the true residual meets tol.  (Keeping the direction
        // here, as the recomputations during the solve do, let the recursion
        // run away once the preconditioner holds inner iterations.)
        if (c->sc_host->active || (c->sc_host->guard & 1) || replacements >= 4 || it >= max_iter) break;
        rc = enqueue_init(c, 1, FIN_REPLACE);
        if (rc) return rc;
        SG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
        SG_HIP(hipStreamSynchronize(s));
        ++replacements;
        if (!c->sc_host->active) break;
    }
    rc = poisson_scatter(c);
    if (rc) return rc;
    SG_HIP(hipEventRecord(c->ev2, s));
    double relres = 0.0, maxrow = 0.0;
    rc = fine_residual_norms(c, &relres, &maxrow);
    if (rc) return rc;
    float ms_all = 0.f, ms_setup = 0.f;
    SG_HIP(hipEventElapsedTime(&ms_all, c->ev0, c->ev2));
    SG_HIP(hipEventElapsedTime(&ms_setup, c->ev0, c->ev1));
    int status = SG_OK;
    if (c->sc_host->guard & 1) {
        status = SG_NAN;
        set_error("poisson solve: NaN in the PCG scalars");
    } else if (c->sc_host->active) {
        status = SG_NOT_CONVERGED;
        set_error("poisson solve: not converged within max_iter");
    }
    if (info) {
        info->iterations = c->sc_host->iters;
        info->relres = relres;
        info->max_row_residual = maxrow;
        info->ms = ms_all;
        info->setup_ms = ms_setup;
        info->status = status;
        info->warm_ms = warm_ms;
        info->warm_relres = warm_relres;
        info->warm_energy = warm_energy;
    }
    return status;
}

}  // namespace sg
