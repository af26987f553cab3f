// Preconditioned conjugate gradients for the projection (stands where
// HYPRE_ParCSRPCGSolve stands, ProjectionP_Mod.f90:57-58), FP64 vectors and
// dot products, multigrid V-cycle preconditioner.  All scalars (alpha, beta,
// norms) stay on the device; the host reads them only to test convergence.
// In mixed precision the cycle is not exactly symmetric, so beta takes the
// Polak-Ribiere (flexible) form -alpha*(z_new,q_old)/(r_old,z_old).
#include "sg_internal.hpp"
#include "sg_reduce.hpp"

namespace sg {

namespace {

constexpr int NT = TX * TY;

__device__ inline double* slot(const Reducer& R, int s) { return R.partials + (size_t)s * NRED * R.max_blocks; }

// r = b - A x (or r = b for a zero start), ||b||^2, ||r||^2
__global__ void __launch_bounds__(256) k_init(Dims D, const double* __restrict__ diag, const double* __restrict__ cE,
                                              const double* __restrict__ cN, const double* __restrict__ b,
                                              double* __restrict__ x, double* __restrict__ r, int warm,
                                              int replace, Reducer R, PcgScalars* sc, double tol) {
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
        r[p] = rp;
        v[0] = bp * bp;
        v[1] = rp * rp;
    }
    double out[2];
    if (grid_reduce<2, NT>(v, slot(R, 0), R.max_blocks, R.tickets + 0, out)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            // replace: the recursive residual met the tolerance; restart from
            // the true residual b - A x if it did not
            if (!replace) {
                sc->bb = out[0];
                sc->tol2 = tol * tol * out[0];
                sc->iters = 0;
            }
            sc->rr = out[1];
            sc->first = 1;
            sc->active = (out[0] > 0.0 && out[1] > sc->tol2) ? 1 : 0;
            sc->rz = sc->rz_old = sc->zq = sc->pq = 0.0;
        }
    }
}

// rz = (r,z), zq = (z,q_old); last block forms beta
__global__ void __launch_bounds__(256) k_dots_rz(Dims D, const double* __restrict__ r, const double* __restrict__ z,
                                                 const double* __restrict__ q, Reducer R, PcgScalars* sc,
                                                 double* beta_out) {
    if (!sc->active) return;
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    double v[2] = {0.0, 0.0};
    if (i < D.ni && j < D.nj) {
        const size_t p = D.off(i, j);
        const double zp = z[p];
        v[0] = r[p] * zp;
        v[1] = zp * q[p];
    }
    double out[2];
    if (grid_reduce<2, NT>(v, slot(R, 1), R.max_blocks, R.tickets + 1, out)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            double beta = 0.0;
            if (!sc->first) {
                if (sc->flexible)
                    beta = -beta_out[1] * out[1] / sc->rz_old;  // beta_out[1] holds alpha
                else
                    beta = out[0] / sc->rz_old;
            }
            sc->rz = out[0];
            sc->zq = out[1];
            beta_out[0] = beta;
        }
    }
}

// p_new = z + beta p_old ; q = A p_new ; pq = (p_new, q); last block: alpha
__global__ void __launch_bounds__(256) k_dir_apply(Dims D, const double* __restrict__ diag,
                                                   const double* __restrict__ cE, const double* __restrict__ cN,
                                                   const double* __restrict__ z, const double* __restrict__ pold,
                                                   double* __restrict__ pnew, double* __restrict__ q, Reducer R,
                                                   PcgScalars* sc, double* ab) {
    if (!sc->active) return;
    const double beta = ab[0];
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    double v[1] = {0.0};
    if (i < D.ni && j < D.nj) {
        const size_t p = D.off(i, j);
        const size_t ld = D.ld;
        auto pn = [&](size_t k) { return z[k] + beta * pold[k]; };
        const double pc = pn(p);
        const double qv = diag[p] * pc - cE[p - 1] * pn(p - 1) - cE[p] * pn(p + 1) - cN[p - ld] * pn(p - ld) -
                          cN[p] * pn(p + ld);
        pnew[p] = pc;
        q[p] = qv;
        v[0] = pc * qv;
    }
    double out[1];
    if (grid_reduce<1, NT>(v, slot(R, 2), R.max_blocks, R.tickets + 2, out)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            sc->pq = out[0];
            ab[1] = sc->rz / out[0];  // alpha
            sc->first = 0;
        }
    }
}

// x += alpha p ; r -= alpha q ; rr = (r,r); last block: convergence test
__global__ void __launch_bounds__(256) k_update(Dims D, const double* __restrict__ p, const double* __restrict__ q,
                                                double* __restrict__ x, double* __restrict__ r, Reducer R,
                                                PcgScalars* sc, const double* ab) {
    if (!sc->active) return;
    const double alpha = ab[1];
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    double v[1] = {0.0};
    if (i < D.ni && j < D.nj) {
        const size_t k = D.off(i, j);
        x[k] += alpha * p[k];
        const double rn = r[k] - alpha * q[k];
        r[k] = rn;
        v[0] = rn * rn;
    }
    double out[1];
    if (grid_reduce<1, NT>(v, slot(R, 3), R.max_blocks, R.tickets + 3, out)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            sc->rr = out[0];
            sc->rz_old = sc->rz;
            sc->iters += 1;
            if (!(out[0] > sc->tol2)) sc->active = 0;
            if (isnan(out[0])) {
                sc->active = 0;
                sc->guard |= 1;
            }
        }
    }
}

// true residual of the fine system: sum of squares and max |.| over fluid rows
__global__ void __launch_bounds__(256) k_true_resid(Dims D, const double* __restrict__ diag,
                                                    const double* __restrict__ cE, const double* __restrict__ cN,
                                                    const double* __restrict__ b, const double* __restrict__ x,
                                                    double* partials, int max_blocks, unsigned int* ticket,
                                                    double* out) {
    __shared__ double ssum[NT / 64], smax[NT / 64];
    __shared__ bool is_last;
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    double s = 0.0, m = 0.0;
    if (i < D.ni && j < D.nj) {
        const size_t p = D.off(i, j);
        const size_t ld = D.ld;
        if (diag[p] > 0.0) {
            const double rr = b[p] - (diag[p] * x[p] - cE[p - 1] * x[p - 1] - cE[p] * x[p + 1] -
                                      cN[p - ld] * x[p - ld] - cN[p] * x[p + ld]);
            s = rr * rr;
            m = fabs(rr);
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        m = fmax(m, __shfl_xor(m, o, 64));
    }
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if ((tid & 63) == 0) {
        ssum[tid >> 6] = s;
        smax[tid >> 6] = m;
    }
    __syncthreads();
    const int nblocks = gridDim.x * gridDim.y;
    const int bid = blockIdx.x + blockIdx.y * gridDim.x;
    if (tid == 0) {
        double a = 0.0, c = 0.0;
        for (int w = 0; w < NT / 64; ++w) {
            a += ssum[w];
            c = fmax(c, smax[w]);
        }
        partials[bid] = a;
        partials[max_blocks + bid] = c;
        __threadfence();
        const unsigned int t = atomicAdd(ticket, 1u);
        is_last = (t == (unsigned int)(nblocks - 1));
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double a = 0.0, c = 0.0;
    for (int bb = tid; bb < nblocks; bb += NT) {
        a += partials[bb];
        c = fmax(c, partials[max_blocks + bb]);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        c = fmax(c, __shfl_xor(c, o, 64));
    }
    if ((tid & 63) == 0) {
        ssum[tid >> 6] = a;
        smax[tid >> 6] = c;
    }
    __syncthreads();
    if (tid == 0) {
        double aa = 0.0, cc = 0.0;
        for (int w = 0; w < NT / 64; ++w) {
            aa += ssum[w];
            cc = fmax(cc, smax[w]);
        }
        out[0] = aa;
        out[1] = cc;
        *ticket = 0u;
        __threadfence();
    }
}

}  // namespace

int fine_residual_norms(Ctx* c, double* relres, double* maxrow) {
    Level& L0 = c->lev[0];
    double* out = c->red.partials + (size_t)7 * NRED * c->red.max_blocks;  // spare slot
    double* parts = c->red.partials + (size_t)5 * NRED * c->red.max_blocks;
    hipLaunchKernelGGL(k_true_resid, grid2d(c->ni, c->nj), block2d(), 0, c->stream, L0.d, L0.A.diag, L0.A.cE,
                       L0.A.cN, c->rhs, c->x, parts, c->red.max_blocks, c->red.tickets + 5, out);
    SG_HIP(hipGetLastError());
    double h[2];
    SG_HIP(hipMemcpyAsync(h, out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    const double bb = c->sc_host->bb;
    *relres = bb > 0.0 ? sqrt(h[0] / bb) : 0.0;
    *maxrow = h[1];
    return SG_OK;
}

int pcg_solve(Ctx* c, sg_solve_info* info) {
    Level& L0 = c->lev[0];
    const Dims D = L0.d;
    hipStream_t s = c->stream;
    const dim3 g = grid2d(c->ni, c->nj), b = block2d();
    double* ab = c->red.partials + (size_t)6 * NRED * c->red.max_blocks;  // beta, alpha
    SG_HIP(hipEventRecord(c->ev0, s));
    if (!c->mg_ready) {
        int rc = mg_setup(c);
        if (rc) return rc;
    }
    SG_HIP(hipEventRecord(c->ev1, s));
    // flexible beta in mixed precision
    const int flexible = c->prm.precision == SG_PREC_MIXED ? 1 : 0;
    SG_HIP(hipMemcpyAsync(&c->sc->flexible, &flexible, sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_init, g, b, 0, s, D, L0.A.diag, L0.A.cE, L0.A.cN, (const double*)c->rhs, c->x, c->r,
                       c->prm.warm_start, 0, c->red, c->sc, c->prm.tol);
    SG_HIP(hipGetLastError());
    SG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    const int max_iter = c->prm.max_iter > 0 ? c->prm.max_iter : 1000;
    const int every = c->prm.check_every > 0 ? c->prm.check_every : 1;
    int it = 0;
    int replacements = 0;
    for (;;) {
    while (c->sc_host->active && it < max_iter) {
        const int nb = (max_iter - it) < every ? (max_iter - it) : every;
        for (int k = 0; k < nb; ++k) {
            double* pold = c->pcur ? c->p1 : c->p0;
            double* pnew = c->pcur ? c->p0 : c->p1;
            int rc = mg_vcycle(c, c->r, c->z);
            if (rc) return rc;
            hipLaunchKernelGGL(k_dots_rz, g, b, 0, s, D, (const double*)c->r, (const double*)c->z,
                               (const double*)c->q, c->red, c->sc, ab);
            hipLaunchKernelGGL(k_dir_apply, g, b, 0, s, D, L0.A.diag, L0.A.cE, L0.A.cN, (const double*)c->z,
                               (const double*)pold, pnew, c->q, c->red, c->sc, ab);
            hipLaunchKernelGGL(k_update, g, b, 0, s, D, (const double*)pnew, (const double*)c->q, c->x, c->r, c->red,
                               c->sc, (const double*)ab);
            SG_HIP(hipGetLastError());
            c->pcur ^= 1;
        }
        it += nb;
        SG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
        SG_HIP(hipStreamSynchronize(s));
    }
    // FP64 residual replacement: the recursive residual drifts from b - A x
    // by rounding; restart CG from the true residual until it meets tol
    if (c->sc_host->active || (c->sc_host->guard & 1) || replacements >= 4 || it >= max_iter) break;
    hipLaunchKernelGGL(k_init, g, b, 0, s, D, L0.A.diag, L0.A.cE, L0.A.cN, (const double*)c->rhs, c->x, c->r, 1, 1,
                       c->red, c->sc, c->prm.tol);
    SG_HIP(hipGetLastError());
    SG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    ++replacements;
    if (!c->sc_host->active) break;
    }
    int rc = poisson_scatter(c);
    if (rc) return rc;
    SG_HIP(hipEventRecord(c->ev2, s));
    double relres = 0.0, maxrow = 0.0;
    rc = fine_residual_norms(c, &relres, &maxrow);
    if (rc) return rc;
    float ms_all = 0.f, ms_setup = 0.f;
    SG_HIP(hipEventElapsedTime(&ms_all, c->ev0, c->ev2));
    SG_HIP(hipEventElapsedTime(&ms_setup, c->ev0, c->ev1));
    int status = SG_OK;
    if (c->sc_host->guard & 1) status = SG_NAN;
    else if (c->sc_host->active) status = SG_NOT_CONVERGED;
    if (info) {
        info->iterations = c->sc_host->iters;
        info->relres = relres;
        info->max_row_residual = maxrow;
        info->ms = ms_all;
        info->setup_ms = ms_setup;
        info->status = status;
    }
    return status;
}

}  // namespace sg
