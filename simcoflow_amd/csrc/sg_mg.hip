// Structured-grid multigrid preconditioner for the cut-cell pressure
// operator (stands where BoomerAMG stands in ProjectionP_Mod.f90:117-133).
//
//  * coarsening: every other cell in x and y (coarse point (I,J) sits on fine
//    cell (2I,2J)); levels down to <= 64 unknowns;
//  * prolongation: operator-dependent (black-box) interpolation -- edge
//    points lump the stencil column/row onto their two coarse neighbours,
//    centre points average their eight interpolated neighbours with the
//    operator's own couplings -- so 1:1000 density jumps and cut-cell
//    apertures shape the coarse basis;
//  * coarse operators: Galerkin P^T A P (9-point, symmetric), built in FP64
//    every solve; FP32 copies drive the mixed-precision cycle;
//  * smoothing (sg_cycle.hip): red-black Gauss-Seidel on every level,
//    reversed order after the coarse-grid correction so the V-cycle is a
//    symmetric preconditioner;
//  * coarsest level: dense inverse in FP64 (Gauss-Jordan in LDS).
#include "sg_internal.hpp"
#include "sg_mg_common.hpp"

namespace sg {

namespace {

// ---- prolongation weights --------------------------------------------------
__global__ void __launch_bounds__(256) k_prol_edges(Dims F, OpV<double> A, Dims C, Prol<double> P) {
    const int I = blockIdx.x * TX + threadIdx.x;
    const int J = blockIdx.y * TY + threadIdx.y;
    if (I >= C.ni || J >= C.nj) return;
    const int fi = 2 * I, fj = 2 * J;
    const size_t cp = C.off(I, J);
    const size_t ld = F.ld;
    P.w[W_INJ][cp] = (A.diag[F.off(fi, fj)] > 0.0) ? 1.0 : 0.0;
    double xW = 0.0, xE = 0.0, yS = 0.0, yN = 0.0;
    if (fi + 1 < F.ni) {
        const size_t f = F.off(fi + 1, fj);
        const double d = A.diag[f];
        if (d > 0.0) {
            const double den = d - A.cN[f] - A.cN[f - ld];
            const double wcol = A.cE[f - 1] + opget(A.cNW, f) + opget(A.cNE, f - ld - 1);
            const double ecol = A.cE[f] + opget(A.cNE, f) + opget(A.cNW, f - ld + 1);
            if (den > 0.0) {
                xW = wcol / den;
                xE = ecol / den;
            }
        }
    }
    if (fj + 1 < F.nj) {
        const size_t f = F.off(fi, fj + 1);
        const double d = A.diag[f];
        if (d > 0.0) {
            const double den = d - A.cE[f] - A.cE[f - 1];
            const double srow = A.cN[f - ld] + opget(A.cNE, f - ld - 1) + opget(A.cNW, f - ld + 1);
            const double nrow = A.cN[f] + opget(A.cNW, f) + opget(A.cNE, f);
            if (den > 0.0) {
                yS = srow / den;
                yN = nrow / den;
            }
        }
    }
    P.w[W_XW][cp] = xW;
    P.w[W_XE][cp] = xE;
    P.w[W_YS][cp] = yS;
    P.w[W_YN][cp] = yN;
}

__global__ void __launch_bounds__(256) k_prol_center(Dims F, OpV<double> A, Dims C, Prol<double> P) {
    const int I = blockIdx.x * TX + threadIdx.x;
    const int J = blockIdx.y * TY + threadIdx.y;
    if (I >= C.ni || J >= C.nj) return;
    const size_t cp = C.off(I, J);
    const int fi = 2 * I + 1, fj = 2 * J + 1;
    double wsw = 0.0, wse = 0.0, wnw = 0.0, wne = 0.0;
    if (fi < F.ni && fj < F.nj) {
        const size_t f = F.off(fi, fj);
        const size_t ld = F.ld;
        const double d = A.diag[f];
        if (d > 0.0) {
            const double cW = A.cE[f - 1], cE = A.cE[f], cS = A.cN[f - ld], cN = A.cN[f];
            const double cSW = opget(A.cNE, f - ld - 1), cSE = opget(A.cNW, f - ld + 1);
            const double cNW = opget(A.cNW, f), cNE = opget(A.cNE, f);
            const size_t c00 = cp, c10 = C.off(I + 1, J), c01 = C.off(I, J + 1), c11 = C.off(I + 1, J + 1);
            double* const* w = P.w;
            wsw = (cW * w[W_YS][c00] + cS * w[W_XW][c00] + cSW * w[W_INJ][c00]) / d;
            wse = (cE * w[W_YS][c10] + cS * w[W_XE][c00] + cSE * w[W_INJ][c10]) / d;
            wnw = (cW * w[W_YN][c00] + cN * w[W_XW][c01] + cNW * w[W_INJ][c01]) / d;
            wne = (cE * w[W_YN][c10] + cN * w[W_XE][c01] + cNE * w[W_INJ][c11]) / d;
        }
    }
    P.w[W_SW][cp] = wsw;
    P.w[W_SE][cp] = wse;
    P.w[W_NW][cp] = wnw;
    P.w[W_NE][cp] = wne;
}

// ---- Galerkin coarse operator ---------------------------------------------
__global__ void __launch_bounds__(256) k_rap(Dims F, OpV<double> A, Dims C, Prol<double> P, Op<double> Ac) {
    const int I = blockIdx.x * TX + threadIdx.x;
    const int J = blockIdx.y * TY + threadIdx.y;
    if (I >= C.ni || J >= C.nj) return;
    // the five coarse neighbours whose coupling this point stores
    const int DI[5] = {I, I + 1, I, I + 1, I - 1};
    const int DJ[5] = {J, J, J + 1, J + 1, J + 1};
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double* const* w = P.w;
    for (int b = -1; b <= 1; ++b) {
        for (int a = -1; a <= 1; ++a) {
            const int fi = 2 * I + a, fj = 2 * J + b;
            if (fi < 0 || fj < 0 || fi >= F.ni || fj >= F.nj) continue;
            const double pfc = pweight<double>(C, w, fi, fj, I, J);
            if (pfc == 0.0) continue;
            for (int dj = -1; dj <= 1; ++dj) {
                for (int di = -1; di <= 1; ++di) {
                    const double afg = aentry<double>(F, A, fi, fj, di, dj);
                    if (afg == 0.0) continue;
                    const int gi = fi + di, gj = fj + dj;
                    const double s = pfc * afg;
#pragma unroll
                    for (int k = 0; k < 5; ++k) acc[k] += s * pweight<double>(C, w, gi, gj, DI[k], DJ[k]);
                }
            }
        }
    }
    const size_t cp = C.off(I, J);
    Ac.diag[cp] = acc[0] > 0.0 ? acc[0] : 0.0;
    Ac.cE[cp] = (I + 1 < C.ni) ? -acc[1] : 0.0;
    Ac.cN[cp] = (J + 1 < C.nj) ? -acc[2] : 0.0;
    Ac.cNE[cp] = (I + 1 < C.ni && J + 1 < C.nj) ? -acc[3] : 0.0;
    Ac.cNW[cp] = (I > 0 && J + 1 < C.nj) ? -acc[4] : 0.0;
}

// FP32 copy of an operator / of prolongation weights (interior only)
__global__ void __launch_bounds__(256) k_to_f32(Dims D, const double* __restrict__ a, float* __restrict__ b) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= D.ni || j >= D.nj) return;
    const size_t p = D.off(i, j);
    b[p] = (float)a[p];
}

// ---- dense inverse of the coarsest operator (one workgroup) ---------------
__global__ void __launch_bounds__(256) k_coarse_inverse(Dims C, OpV<double> A, double* inv, float* invf) {
    __shared__ double a[COARSEST_MAX][COARSEST_MAX];
    __shared__ double m[COARSEST_MAX][COARSEST_MAX];
    __shared__ double amax_s;
    const int n = C.ni * C.nj;
    const int tid = threadIdx.x;
    for (int e = tid; e < n * n; e += blockDim.x) {
        const int r = e / n, q = e % n;
        const int ri = r % C.ni, rj = r / C.ni, qi = q % C.ni, qj = q / C.ni;
        const int di = qi - ri, dj = qj - rj;
        double v = 0.0;
        if (di >= -1 && di <= 1 && dj >= -1 && dj <= 1) v = aentry<double>(C, A, ri, rj, di, dj);
        a[r][q] = v;
        m[r][q] = (r == q) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (tid == 0) {
        double mx = 0.0;
        for (int k = 0; k < n; ++k) mx = fmax(mx, fabs(a[k][k]));
        amax_s = mx;
    }
    __syncthreads();
    const double amax = amax_s;
    for (int k = 0; k < n; ++k) {
        const double piv = a[k][k];
        __syncthreads();
        if (!(fabs(piv) > 1e-13 * amax)) {
            for (int e = tid; e < n; e += blockDim.x) {
                a[k][e] = 0.0;
                a[e][k] = 0.0;
                m[k][e] = 0.0;
                m[e][k] = 0.0;
            }
            __syncthreads();
            continue;
        }
        const double ip = 1.0 / piv;
        for (int e = tid; e < n; e += blockDim.x) {
            a[k][e] *= ip;
            m[k][e] *= ip;
        }
        __syncthreads();
        for (int e = tid; e < n * n; e += blockDim.x) {
            const int r = e / n, q = e % n;
            if (r == k) continue;
            const double f = a[r][k];
            if (f != 0.0) {
                a[r][q] -= (q == k) ? 0.0 : f * a[k][q];
                m[r][q] -= f * m[k][q];
            }
        }
        __syncthreads();
        for (int r = tid; r < n; r += blockDim.x)
            if (r != k) a[r][k] = 0.0;
        __syncthreads();
    }
    for (int e = tid; e < n * n; e += blockDim.x) {
        const int r = e / n, q = e % n;
        inv[e] = m[r][q];
        invf[e] = (float)m[r][q];
    }
}

}  // namespace

int mg_allocate(Ctx* c) {
    int l = 0;
    c->lev[0].d = make_dims(c->ni, c->nj);
    while ((size_t)c->lev[l].d.ni * c->lev[l].d.nj > (size_t)COARSEST_MAX && l + 1 < MAX_LEVELS) {
        const Dims& f = c->lev[l].d;
        c->lev[l + 1].d = make_dims((f.ni + 1) / 2, (f.nj + 1) / 2);
        ++l;
    }
    c->nlev = l + 1;
    for (int k = 0; k < c->nlev; ++k) {
        Level& L = c->lev[k];
        const size_t n = L.d.alloc_elems();
        auto d = [&](void) { return (double*)dev_alloc(c, n * sizeof(double)); };
        auto f = [&](void) { return (float*)dev_alloc(c, n * sizeof(float)); };
        L.A.diag = d();
        L.A.cE = d();
        L.A.cN = d();
        L.Af.diag = f();
        L.Af.cE = f();
        L.Af.cN = f();
        if (k > 0) {
            L.A.cNE = d();
            L.A.cNW = d();
            L.Af.cNE = f();
            L.Af.cNW = f();
        }
        L.v.x = d();
        L.v.b = d();
        L.v.r = d();
        L.vf.x = f();
        L.vf.b = f();
        L.vf.r = f();
        if (!L.A.diag || !L.A.cE || !L.A.cN || !L.v.x || !L.v.b || !L.v.r || !L.vf.x || !L.vf.b || !L.vf.r)
            return SG_ERR_ALLOC;
    }
    // prolongation weights live at the coarse resolution of the next level
    for (int k = 0; k + 1 < c->nlev; ++k) {
        const size_t n = c->lev[k + 1].d.alloc_elems();
        for (int w = 0; w < 9; ++w) {
            c->lev[k].P.w[w] = (double*)dev_alloc(c, n * sizeof(double));
            c->lev[k].Pf.w[w] = (float*)dev_alloc(c, n * sizeof(float));
            if (!c->lev[k].P.w[w] || !c->lev[k].Pf.w[w]) return SG_ERR_ALLOC;
        }
    }
    const Dims& C = c->lev[c->nlev - 1].d;
    c->coarse_n = C.ni * C.nj;
    c->coarse_inv = (double*)dev_alloc(c, sizeof(double) * COARSEST_MAX * COARSEST_MAX);
    c->coarse_inv_f = (float*)dev_alloc(c, sizeof(float) * COARSEST_MAX * COARSEST_MAX);
    if (!c->coarse_inv || !c->coarse_inv_f) return SG_ERR_ALLOC;
    return SG_OK;
}

int mg_setup(Ctx* c) {
    hipStream_t s = c->stream;
    const bool mixed = c->prm.precision == SG_PREC_MIXED;
    for (int l = 0; l + 1 < c->nlev; ++l) {
        Level& F = c->lev[l];
        Level& C = c->lev[l + 1];
        OpV<double> A = opv(F.A);
        hipLaunchKernelGGL(k_prol_edges, grid2d(C.d.ni, C.d.nj), block2d(), 0, s, F.d, A, C.d, F.P);
        hipLaunchKernelGGL(k_prol_center, grid2d(C.d.ni, C.d.nj), block2d(), 0, s, F.d, A, C.d, F.P);
        hipLaunchKernelGGL(k_rap, grid2d(C.d.ni, C.d.nj), block2d(), 0, s, F.d, A, C.d, F.P, C.A);
        SG_HIP(hipGetLastError());
        if (mixed) {
            for (int w = 0; w < 9; ++w)
                hipLaunchKernelGGL(k_to_f32, grid2d(C.d.ni, C.d.nj), block2d(), 0, s, C.d, F.P.w[w], F.Pf.w[w]);
            const double* src[5] = {C.A.diag, C.A.cE, C.A.cN, C.A.cNE, C.A.cNW};
            float* dst[5] = {C.Af.diag, C.Af.cE, C.Af.cN, C.Af.cNE, C.Af.cNW};
            for (int k = 0; k < 5; ++k)
                hipLaunchKernelGGL(k_to_f32, grid2d(C.d.ni, C.d.nj), block2d(), 0, s, C.d, src[k], dst[k]);
            SG_HIP(hipGetLastError());
        }
    }
    Level& Lc = c->lev[c->nlev - 1];
    hipLaunchKernelGGL(k_coarse_inverse, dim3(1), dim3(256), 0, s, Lc.d, opv(Lc.A), c->coarse_inv, c->coarse_inv_f);
    SG_HIP(hipGetLastError());
    c->mg_ready = 1;
    return SG_OK;
}

}  // namespace sg
