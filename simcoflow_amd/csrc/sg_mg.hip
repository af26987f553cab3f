// Structured-grid multigrid preconditioner for the cut-cell pressure
// operator (stands where BoomerAMG stands in ProjectionP_Mod.f90:117-133).
//
//  * coarsening: every other cell in x and y (coarse point (I,J) sits on fine
//    cell (2I,2J)); levels down to <= 64 unknowns.  On grids of cells at
//    least 1.4x wider than tall the first coarsening is in y only (coarse
//    (I,J) on fine (I,2J)): the N couplings there are twice the E ones or
//    more, and red-black point smoothing reduces the error along the strong
//    direction only -- coarsening that direction alone leaves an isotropic
//    level 1 (semi-coarsening);
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
__global__ void __launch_bounds__(256) k_prol_edges(Dims F, OpV<double> A, Dims C, Prol<double> P, Prol<float> Pf) {
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
    if (Pf.w[0]) {
        Pf.w[W_INJ][cp] = (float)P.w[W_INJ][cp];
        Pf.w[W_XW][cp] = (float)xW;
        Pf.w[W_XE][cp] = (float)xE;
        Pf.w[W_YS][cp] = (float)yS;
        Pf.w[W_YN][cp] = (float)yN;
    }
}

__global__ void __launch_bounds__(256) k_prol_center(Dims F, OpV<double> A, Dims C, Prol<double> P, Prol<float> Pf) {
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
    if (Pf.w[0]) {
        Pf.w[W_SW][cp] = (float)wsw;
        Pf.w[W_SE][cp] = (float)wse;
        Pf.w[W_NW][cp] = (float)wnw;
        Pf.w[W_NE][cp] = (float)wne;
    }
}

// y-only coarsening of the 5-point fine level: coarse (I,J) on fine (I,2J)
// (injection); the odd row (I,2J+1) from (I,J) and (I,J+1) with the row's x
// couplings lumped onto its diagonal, as the edge points above
__global__ void __launch_bounds__(256) k_prol_semi(Dims F, OpV<double> A, Dims C, Prol<double> P, Prol<float> Pf) {
    const int I = blockIdx.x * TX + threadIdx.x;
    const int J = blockIdx.y * TY + threadIdx.y;
    if (I >= C.ni || J >= C.nj) return;
    const size_t cp = C.off(I, J);
    const size_t ld = F.ld;
    const double inj = (A.diag[F.off(I, 2 * J)] > 0.0) ? 1.0 : 0.0;
    double yS = 0.0, yN = 0.0;
    if (2 * J + 1 < F.nj) {
        const size_t f = F.off(I, 2 * J + 1);
        const double d = A.diag[f];
        if (d > 0.0) {
            const double den = d - A.cE[f] - A.cE[f - 1];
            if (den > 0.0) {
                yS = A.cN[f - ld] / den;
                yN = A.cN[f] / den;
            }
        }
    }
    const double w[9] = {inj, 0.0, 0.0, yS, yN, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 9; ++k) P.w[k][cp] = w[k];
    if (Pf.w[0]) {
#pragma unroll
        for (int k = 0; k < 9; ++k) Pf.w[k][cp] = (float)w[k];
    }
}

// ---- Galerkin coarse operator ---------------------------------------------
// Ac(C1,C2) = sum_f sum_g P(f,C1) A(f,g) P(g,C2) for C1 = (I,J) and the five
// stored neighbours C2 - C1 in {(0,0),(1,0),(0,1),(1,1),(-1,1)}.  The fine
// cells f with P(f,C1) != 0 are (2I+a, 2J+b), |a|,|b| <= 1; g = f + (di,dj)
// runs over the stencil of f.  Every offset is a compile-time constant once
// the loops are unrolled, so the parity cases of the prolongation and the
// target each product lands in are resolved statically: the kernel is a
// straight sum of products of coalesced loads.  Ring cells of the operator
// and of the weights are zero, so no bounds tests are needed.
__device__ __forceinline__ int rap_target(int dI, int dJ) {
    return (dI == 0 && dJ == 0) ? 0 : (dI == 1 && dJ == 0) ? 1 : (dI == 0 && dJ == 1) ? 2
         : (dI == 1 && dJ == 1) ? 3 : (dI == -1 && dJ == 1) ? 4 : -1;
}

template <bool NINE>
__global__ void __launch_bounds__(256) k_rap(Dims F, OpV<double> A, Dims C, Prol<double> P, Op<double> Ac, Op<float> Af) {
    const int I = blockIdx.x * TX + threadIdx.x;
    const int J = blockIdx.y * TY + threadIdx.y;
    if (I >= C.ni || J >= C.nj) return;
    const size_t ldf = F.ld, ldc = C.ld;
    const size_t c = C.off(I, J);
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // P(g, parent) for g = (2I+ga, 2J+gb): add s * weight to the targets
    auto scatter = [&](double s, int ga, int gb) {
        const bool ox = ga & 1, oy = gb & 1;
        if (!ox && !oy) {
            const int pI = ga / 2, pJ = gb / 2, k = rap_target(pI, pJ);
            if (k >= 0) acc[k] += s * P.w[W_INJ][c + (long)pJ * (long)ldc + pI];
        } else if (ox && !oy) {
            const int I0 = (ga - 1) / 2, J0 = gb / 2;
            const size_t q = c + (long)J0 * (long)ldc + I0;
            const int k0 = rap_target(I0, J0), k1 = rap_target(I0 + 1, J0);
            if (k0 >= 0) acc[k0] += s * P.w[W_XW][q];
            if (k1 >= 0) acc[k1] += s * P.w[W_XE][q];
        } else if (!ox && oy) {
            const int I0 = ga / 2, J0 = (gb - 1) / 2;
            const size_t q = c + (long)J0 * (long)ldc + I0;
            const int k0 = rap_target(I0, J0), k1 = rap_target(I0, J0 + 1);
            if (k0 >= 0) acc[k0] += s * P.w[W_YS][q];
            if (k1 >= 0) acc[k1] += s * P.w[W_YN][q];
        } else {
            const int I0 = (ga - 1) / 2, J0 = (gb - 1) / 2;
            const size_t q = c + (long)J0 * (long)ldc + I0;
            const int k0 = rap_target(I0, J0), k1 = rap_target(I0 + 1, J0);
            const int k2 = rap_target(I0, J0 + 1), k3 = rap_target(I0 + 1, J0 + 1);
            if (k0 >= 0) acc[k0] += s * P.w[W_SW][q];
            if (k1 >= 0) acc[k1] += s * P.w[W_SE][q];
            if (k2 >= 0) acc[k2] += s * P.w[W_NW][q];
            if (k3 >= 0) acc[k3] += s * P.w[W_NE][q];
        }
    };
#pragma unroll
    for (int b = -1; b <= 1; ++b) {
#pragma unroll
        for (int a = -1; a <= 1; ++a) {
            // P(f, (I,J)) for f = (2I+a, 2J+b)
            double wf;
            if (a == 0 && b == 0) wf = P.w[W_INJ][c];
            else if (b == 0) wf = a > 0 ? P.w[W_XW][c] : P.w[W_XE][c - 1];
            else if (a == 0) wf = b > 0 ? P.w[W_YS][c] : P.w[W_YN][c - ldc];
            else if (a > 0) wf = b > 0 ? P.w[W_SW][c] : P.w[W_NW][c - ldc];
            else wf = b > 0 ? P.w[W_SE][c - 1] : P.w[W_NE][c - ldc - 1];
            const size_t f = F.off(2 * I + a, 2 * J + b);
            scatter(wf * A.diag[f], a, b);
            scatter(wf * -A.cE[f], a + 1, b);
            scatter(wf * -A.cE[f - 1], a - 1, b);
            scatter(wf * -A.cN[f], a, b + 1);
            scatter(wf * -A.cN[f - ldf], a, b - 1);
            if (NINE) {
                scatter(wf * -A.cNE[f], a + 1, b + 1);
                scatter(wf * -A.cNE[f - ldf - 1], a - 1, b - 1);
                scatter(wf * -A.cNW[f], a - 1, b + 1);
                scatter(wf * -A.cNW[f - ldf + 1], a + 1, b - 1);
            }
        }
    }
    const double d = acc[0] > 0.0 ? acc[0] : 0.0;
    const double e = (I + 1 < C.ni) ? -acc[1] : 0.0;
    const double n = (J + 1 < C.nj) ? -acc[2] : 0.0;
    const double ne = (I + 1 < C.ni && J + 1 < C.nj) ? -acc[3] : 0.0;
    const double nw = (I > 0 && J + 1 < C.nj) ? -acc[4] : 0.0;
    Ac.diag[c] = d;
    Ac.cE[c] = e;
    Ac.cN[c] = n;
    Ac.cNE[c] = ne;
    Ac.cNW[c] = nw;
    if (Af.diag) {
        Af.diag[c] = (float)d;
        Af.cE[c] = (float)e;
        Af.cN[c] = (float)n;
        Af.cNE[c] = (float)ne;
        Af.cNW[c] = (float)nw;
    }
}

// Galerkin operator of the y-coarsened level 1 from the 5-point level 0:
// the fine cells carrying coarse (I,J) are (I,2J+b), b in {-1,0,1}; a fine
// cell g = (I+ga, 2J+gb) belongs to coarse (I+ga, J+gb/2) by injection (gb
// even) or to (I+ga, J0) and (I+ga, J0+1), J0 = J+(gb-1)/2, by the odd-row
// weights (gb odd).  The result is 9-point, stored like k_rap's.
__global__ void __launch_bounds__(256) k_rap_semi(Dims F, OpV<double> A, Dims C, Prol<double> P, Op<double> Ac,
                                                  Op<float> Af) {
    const int I = blockIdx.x * TX + threadIdx.x;
    const int J = blockIdx.y * TY + threadIdx.y;
    if (I >= C.ni || J >= C.nj) return;
    const size_t ldf = F.ld, ldc = C.ld;
    const size_t c = C.off(I, J);
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    auto scatter = [&](double s, int ga, int gb) {
        if ((gb & 1) == 0) {
            const int k = rap_target(ga, gb / 2);
            if (k >= 0) acc[k] += s * P.w[W_INJ][c + (long)(gb / 2) * (long)ldc + ga];
        } else {
            const int J0 = (gb - 1) / 2;  // exact: gb - 1 is even
            const size_t q = c + (long)J0 * (long)ldc + ga;
            const int k0 = rap_target(ga, J0), k1 = rap_target(ga, J0 + 1);
            if (k0 >= 0) acc[k0] += s * P.w[W_YS][q];
            if (k1 >= 0) acc[k1] += s * P.w[W_YN][q];
        }
    };
#pragma unroll
    for (int b = -1; b <= 1; ++b) {
        const double wf = b == 0 ? P.w[W_INJ][c] : (b > 0 ? P.w[W_YS][c] : P.w[W_YN][c - ldc]);
        const size_t f = F.off(I, 2 * J + b);
        scatter(wf * A.diag[f], 0, b);
        scatter(wf * -A.cE[f], 1, b);
        scatter(wf * -A.cE[f - 1], -1, b);
        scatter(wf * -A.cN[f], 0, b + 1);
        scatter(wf * -A.cN[f - ldf], 0, b - 1);
    }
    const double d = acc[0] > 0.0 ? acc[0] : 0.0;
    const double e = (I + 1 < C.ni) ? -acc[1] : 0.0;
    const double n = (J + 1 < C.nj) ? -acc[2] : 0.0;
    const double ne = (I + 1 < C.ni && J + 1 < C.nj) ? -acc[3] : 0.0;
    const double nw = (I > 0 && J + 1 < C.nj) ? -acc[4] : 0.0;
    Ac.diag[c] = d;
    Ac.cE[c] = e;
    Ac.cN[c] = n;
    Ac.cNE[c] = ne;
    Ac.cNW[c] = nw;
    if (Af.diag) {
        Af.diag[c] = (float)d;
        Af.cE[c] = (float)e;
        Af.cN[c] = (float)n;
        Af.cNE[c] = (float)ne;
        Af.cNW[c] = (float)nw;
    }
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
        if (l == 0 && c->semi0)
            c->lev[1].d = make_dims(f.ni, (f.nj + 1) / 2);
        else
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
        const Prol<float> Pf = mixed ? F.Pf : Prol<float>{};
        const Op<float> Af = mixed ? C.Af : Op<float>{};
        if (l == 0 && c->semi0) {
            hipLaunchKernelGGL(k_prol_semi, grid2d(C.d.ni, C.d.nj), block2d(), 0, s, F.d, A, C.d, F.P, Pf);
            hipLaunchKernelGGL(k_rap_semi, grid2d(C.d.ni, C.d.nj), block2d(), 0, s, F.d, A, C.d, F.P, C.A, Af);
            SG_HIP(hipGetLastError());
            continue;
        }
        hipLaunchKernelGGL(k_prol_edges, grid2d(C.d.ni, C.d.nj), block2d(), 0, s, F.d, A, C.d, F.P, Pf);
        hipLaunchKernelGGL(k_prol_center, grid2d(C.d.ni, C.d.nj), block2d(), 0, s, F.d, A, C.d, F.P, Pf);
        if (l == 0)
            hipLaunchKernelGGL(k_rap<false>, grid2d(C.d.ni, C.d.nj), block2d(), 0, s, F.d, A, C.d, F.P, C.A, Af);
        else
            hipLaunchKernelGGL(k_rap<true>, grid2d(C.d.ni, C.d.nj), block2d(), 0, s, F.d, A, C.d, F.P, C.A, Af);
        SG_HIP(hipGetLastError());
    }
    Level& Lc = c->lev[c->nlev - 1];
    hipLaunchKernelGGL(k_coarse_inverse, dim3(1), dim3(256), 0, s, Lc.d, opv(Lc.A), c->coarse_inv, c->coarse_inv_f);
    SG_HIP(hipGetLastError());
    c->mg_ready = 1;
    return SG_OK;
}

}  // namespace sg
