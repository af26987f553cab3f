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
//  * smoothing: red-black Gauss-Seidel on the 5-point fine level, four-colour
//    Gauss-Seidel on the 9-point coarse levels, reversed order after the
//    coarse-grid correction so the V-cycle is a symmetric preconditioner;
//  * coarsest level: dense inverse in FP64 (Gauss-Jordan in LDS).
#include "sg_internal.hpp"

namespace sg {

namespace {

template <typename T>
struct OpV {
    const T *diag, *cE, *cN, *cNE, *cNW;  // cNE/cNW null on the 5-point level
};

template <typename T>
__device__ inline T opget(const T* a, size_t p) {
    return a ? a[p] : T(0);
}

// prolongation weight P(g, D) for fine cell g and coarse point D
template <typename T>
__device__ inline T pweight(const Dims& C, T* const* w, int gi, int gj, int DI, int DJ) {
    const int pi = gi & 1, pj = gj & 1;
    if (!pi && !pj) {
        return (gi >> 1) == DI && (gj >> 1) == DJ ? w[W_INJ][C.off(DI, DJ)] : T(0);
    }
    if (pi && !pj) {
        const int I0 = (gi - 1) >> 1, J0 = gj >> 1;
        if (DJ != J0) return T(0);
        if (DI == I0) return w[W_XW][C.off(I0, J0)];
        if (DI == I0 + 1) return w[W_XE][C.off(I0, J0)];
        return T(0);
    }
    if (!pi && pj) {
        const int I0 = gi >> 1, J0 = (gj - 1) >> 1;
        if (DI != I0) return T(0);
        if (DJ == J0) return w[W_YS][C.off(I0, J0)];
        if (DJ == J0 + 1) return w[W_YN][C.off(I0, J0)];
        return T(0);
    }
    const int I0 = (gi - 1) >> 1, J0 = (gj - 1) >> 1;
    const int a = DI - I0, b = DJ - J0;
    if (a < 0 || a > 1 || b < 0 || b > 1) return T(0);
    const int k = W_SW + a + 2 * b;  // SW, SE, NW, NE
    return w[k][C.off(I0, J0)];
}

// coupling A(f, f+(di,dj)) as a matrix entry (negative of the stored c)
template <typename T>
__device__ inline T aentry(const Dims& F, const OpV<T>& A, int i, int j, int di, int dj) {
    const size_t p = F.off(i, j);
    const size_t ld = F.ld;
    if (di == 0 && dj == 0) return A.diag[p];
    if (dj == 0) return di > 0 ? -A.cE[p] : -A.cE[p - 1];
    if (di == 0) return dj > 0 ? -A.cN[p] : -A.cN[p - ld];
    if (!A.cNE) return T(0);
    if (di > 0 && dj > 0) return -A.cNE[p];
    if (di < 0 && dj < 0) return -A.cNE[p - ld - 1];
    if (di < 0 && dj > 0) return -A.cNW[p];
    return -A.cNW[p - ld + 1];  // (di > 0 && dj < 0)
}

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

// ---- cycle kernels ----------------------------------------------------------
// red-black Gauss-Seidel half sweep on the 5-point level; FIRST: the initial
// guess is zero (colour 0 takes b/diag, colour 1 is cleared)
template <typename T, typename TB, bool FIRST>
__global__ void __launch_bounds__(256) k_rb5(Dims D, OpV<T> A, const TB* __restrict__ b, T* __restrict__ x, int col) {
    const int t = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int i = 2 * t + ((j + col) & 1);
    if (FIRST) {
        // cover both colours of this pair of cells
        const int i0 = 2 * t, i1 = 2 * t + 1;
        if (j >= D.nj) return;
        if (i0 < D.ni) {
            const size_t p = D.off(i0, j);
            const bool mine = ((i0 + j) & 1) == col;
            const T d = A.diag[p];
            x[p] = (mine && d > T(0)) ? (T)b[p] / d : T(0);
        }
        if (i1 < D.ni) {
            const size_t p = D.off(i1, j);
            const bool mine = ((i1 + j) & 1) == col;
            const T d = A.diag[p];
            x[p] = (mine && d > T(0)) ? (T)b[p] / d : T(0);
        }
        return;
    }
    if (i >= D.ni || j >= D.nj) return;
    const size_t p = D.off(i, j);
    const size_t ld = D.ld;
    const T d = A.diag[p];
    T s = (T)b[p] + A.cE[p - 1] * x[p - 1] + A.cE[p] * x[p + 1] + A.cN[p - ld] * x[p - ld] + A.cN[p] * x[p + ld];
    x[p] = d > T(0) ? s / d : T(0);
}

// four-colour Gauss-Seidel on a 9-point level, colour = (i&1) + 2*(j&1)
template <typename T, bool FIRST>
__global__ void __launch_bounds__(256) k_gs9(Dims D, OpV<T> A, const T* __restrict__ b, T* __restrict__ x, int col) {
    const int ci = col & 1, cj = col >> 1;
    const int i = 2 * (blockIdx.x * TX + threadIdx.x) + ci;
    const int j = 2 * (blockIdx.y * TY + threadIdx.y) + cj;
    if (FIRST) {
        // zero initial guess: this colour takes b/diag, the other three are cleared
        const int i0 = i - ci, j0 = j - cj;
        for (int bj = 0; bj < 2; ++bj)
            for (int bi = 0; bi < 2; ++bi) {
                const int ii = i0 + bi, jj = j0 + bj;
                if (ii >= D.ni || jj >= D.nj) continue;
                const size_t p = D.off(ii, jj);
                const T d = A.diag[p];
                x[p] = (bi == ci && bj == cj && d > T(0)) ? b[p] / d : T(0);
            }
        return;
    }
    if (i >= D.ni || j >= D.nj) return;
    const size_t p = D.off(i, j);
    const size_t ld = D.ld;
    const T d = A.diag[p];
    T s = b[p] + A.cE[p - 1] * x[p - 1] + A.cE[p] * x[p + 1] + A.cN[p - ld] * x[p - ld] + A.cN[p] * x[p + ld] +
          A.cNE[p] * x[p + ld + 1] + A.cNE[p - ld - 1] * x[p - ld - 1] + A.cNW[p] * x[p + ld - 1] +
          A.cNW[p - ld + 1] * x[p - ld + 1];
    x[p] = d > T(0) ? s / d : T(0);
}

// r = b - A x
template <typename T, typename TB, bool NINE>
__global__ void __launch_bounds__(256) k_resid(Dims D, OpV<T> A, const TB* __restrict__ b, const T* __restrict__ x,
                                               T* __restrict__ r) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= D.ni || j >= D.nj) return;
    const size_t p = D.off(i, j);
    const size_t ld = D.ld;
    T s = (T)b[p] - A.diag[p] * x[p] + A.cE[p - 1] * x[p - 1] + A.cE[p] * x[p + 1] + A.cN[p - ld] * x[p - ld] +
          A.cN[p] * x[p + ld];
    if (NINE) {
        s += A.cNE[p] * x[p + ld + 1] + A.cNE[p - ld - 1] * x[p - ld - 1] + A.cNW[p] * x[p + ld - 1] +
             A.cNW[p - ld + 1] * x[p - ld + 1];
    }
    r[p] = s;
}

// coarse b = P^T r
template <typename T>
__global__ void __launch_bounds__(256) k_restrict(Dims F, Dims C, Prol<T> P, const T* __restrict__ r, T* __restrict__ bc) {
    const int I = blockIdx.x * TX + threadIdx.x;
    const int J = blockIdx.y * TY + threadIdx.y;
    if (I >= C.ni || J >= C.nj) return;
    const int fi = 2 * I, fj = 2 * J;
    const size_t ld = F.ld;
    const size_t f = F.off(fi, fj);
    T* const* w = P.w;
    const size_t c = C.off(I, J), cw = C.off(I - 1, J), cs = C.off(I, J - 1), csw = C.off(I - 1, J - 1);
    T s = w[W_INJ][c] * r[f];
    s += w[W_XW][c] * r[f + 1] + w[W_XE][cw] * r[f - 1];
    s += w[W_YS][c] * r[f + ld] + w[W_YN][cs] * r[f - ld];
    s += w[W_SW][c] * r[f + ld + 1] + w[W_SE][cw] * r[f + ld - 1] + w[W_NW][cs] * r[f - ld + 1] +
         w[W_NE][csw] * r[f - ld - 1];
    bc[c] = s;
}

// x += P e (fine cells)
template <typename T>
__global__ void __launch_bounds__(256) k_prolong(Dims F, Dims C, Prol<T> P, const T* __restrict__ e, T* __restrict__ x) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= F.ni || j >= F.nj) return;
    T* const* w = P.w;
    const int pi = i & 1, pj = j & 1;
    const int I = i >> 1, J = j >> 1;
    const size_t c = C.off(I, J);
    T v;
    if (!pi && !pj) {
        v = w[W_INJ][c] * e[c];
    } else if (pi && !pj) {
        v = w[W_XW][c] * e[c] + w[W_XE][c] * e[c + 1];
    } else if (!pi && pj) {
        v = w[W_YS][c] * e[c] + w[W_YN][c] * e[c + C.ld];
    } else {
        v = w[W_SW][c] * e[c] + w[W_SE][c] * e[c + 1] + w[W_NW][c] * e[c + C.ld] + w[W_NE][c] * e[c + C.ld + 1];
    }
    x[F.off(i, j)] += v;
}

template <typename T, typename TB>
__global__ void __launch_bounds__(256) k_coarse_solve(Dims C, const T* __restrict__ inv, const TB* __restrict__ b,
                                                      T* __restrict__ x) {
    const int n = C.ni * C.nj;
    const int r = threadIdx.x;
    if (r >= n) return;
    T s = T(0);
    for (int q = 0; q < n; ++q) s += inv[(size_t)r * n + q] * (T)b[C.off(q % C.ni, q / C.ni)];
    x[C.off(r % C.ni, r / C.ni)] = s;
}

template <typename TD, typename TS>
__global__ void __launch_bounds__(256) k_convert(Dims D, const TS* __restrict__ a, TD* __restrict__ b) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= D.ni || j >= D.nj) return;
    const size_t p = D.off(i, j);
    b[p] = (TD)a[p];
}

template <typename T>
OpV<T> opv(const Op<T>& o) {
    return OpV<T>{o.diag, o.cE, o.cN, o.cNE, o.cNW};
}

inline dim3 grid_half(int ni, int nj) { return dim3(((ni + 1) / 2 + TX - 1) / TX, (nj + TY - 1) / TY); }
inline dim3 grid_quarter(int ni, int nj) { return dim3(((ni + 1) / 2 + TX - 1) / TX, ((nj + 1) / 2 + TY - 1) / TY); }

template <typename T>
struct LvAccess;
template <>
struct LvAccess<double> {
    static Op<double>& op(Level& L) { return L.A; }
    static Prol<double>& pr(Level& L) { return L.P; }
    static LevelVecs<double>& vec(Level& L) { return L.v; }
    static double* cinv(Ctx* c) { return c->coarse_inv; }
};
template <>
struct LvAccess<float> {
    static Op<float>& op(Level& L) { return L.Af; }
    static Prol<float>& pr(Level& L) { return L.Pf; }
    static LevelVecs<float>& vec(Level& L) { return L.vf; }
    static float* cinv(Ctx* c) { return c->coarse_inv_f; }
};

// smoothing sweeps on level l; FIRST_ZERO: x starts at zero
template <typename T, typename TB>
void smooth(Ctx* c, int l, const TB* b, T* x, int sweeps, bool first_zero, bool reverse) {
    Level& L = c->lev[l];
    const Dims D = L.d;
    OpV<T> A = opv(LvAccess<T>::op(L));
    hipStream_t s = c->stream;
    for (int sw = 0; sw < sweeps; ++sw) {
        const bool z = first_zero && sw == 0;
        if (l == 0) {
            const int c0 = reverse ? 1 : 0, c1 = reverse ? 0 : 1;
            if (z)
                hipLaunchKernelGGL((k_rb5<T, TB, true>), grid_half(D.ni, D.nj), block2d(), 0, s, D, A, b, x, c0);
            else
                hipLaunchKernelGGL((k_rb5<T, TB, false>), grid_half(D.ni, D.nj), block2d(), 0, s, D, A, b, x, c0);
            hipLaunchKernelGGL((k_rb5<T, TB, false>), grid_half(D.ni, D.nj), block2d(), 0, s, D, A, b, x, c1);
        } else {
            for (int k = 0; k < 4; ++k) {
                const int col = reverse ? 3 - k : k;
                if (z && k == 0)
                    hipLaunchKernelGGL((k_gs9<T, true>), grid_quarter(D.ni, D.nj), block2d(), 0, s, D, A,
                                       (const T*)b, x, col);
                else
                    hipLaunchKernelGGL((k_gs9<T, false>), grid_quarter(D.ni, D.nj), block2d(), 0, s, D, A,
                                       (const T*)b, x, col);
            }
        }
    }
}

template <typename T, typename TB>
int vcycle_level(Ctx* c, int l, const TB* b, T* x) {
    Level& L = c->lev[l];
    hipStream_t s = c->stream;
    if (l == c->nlev - 1) {
        hipLaunchKernelGGL((k_coarse_solve<T, TB>), dim3(1), dim3(256), 0, s, L.d, LvAccess<T>::cinv(c), b, x);
        SG_HIP(hipGetLastError());
        return SG_OK;
    }
    Level& Lc = c->lev[l + 1];
    const int pre = c->prm.mg_pre > 0 ? c->prm.mg_pre : 1;
    const int post = c->prm.mg_post > 0 ? c->prm.mg_post : 1;
    smooth<T, TB>(c, l, b, x, pre, true, false);
    T* r = LvAccess<T>::vec(L).r;
    OpV<T> A = opv(LvAccess<T>::op(L));
    if (l == 0)
        hipLaunchKernelGGL((k_resid<T, TB, false>), grid2d(L.d.ni, L.d.nj), block2d(), 0, s, L.d, A, b, x, r);
    else
        hipLaunchKernelGGL((k_resid<T, TB, true>), grid2d(L.d.ni, L.d.nj), block2d(), 0, s, L.d, A, b, x, r);
    LevelVecs<T>& vc = LvAccess<T>::vec(Lc);
    hipLaunchKernelGGL((k_restrict<T>), grid2d(Lc.d.ni, Lc.d.nj), block2d(), 0, s, L.d, Lc.d, LvAccess<T>::pr(L),
                       (const T*)r, vc.b);
    int rc = vcycle_level<T, T>(c, l + 1, vc.b, vc.x);
    if (rc) return rc;
    hipLaunchKernelGGL((k_prolong<T>), grid2d(L.d.ni, L.d.nj), block2d(), 0, s, L.d, Lc.d, LvAccess<T>::pr(L),
                       (const T*)vc.x, x);
    smooth<T, TB>(c, l, b, x, post, false, true);
    SG_HIP(hipGetLastError());
    return SG_OK;
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

// z = M^{-1} r on the fine level (padded FP64 vectors)
int mg_vcycle(Ctx* c, const double* r, double* z) {
    if (c->prm.precision == SG_PREC_MIXED) {
        Level& L0 = c->lev[0];
        int rc = vcycle_level<float, double>(c, 0, r, L0.vf.x);
        if (rc) return rc;
        hipLaunchKernelGGL((k_convert<double, float>), grid2d(L0.d.ni, L0.d.nj), block2d(), 0, c->stream, L0.d,
                           (const float*)L0.vf.x, z);
        SG_HIP(hipGetLastError());
        return SG_OK;
    }
    return vcycle_level<double, double>(c, 0, r, z);
}

}  // namespace sg
