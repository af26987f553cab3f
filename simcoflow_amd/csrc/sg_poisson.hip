// Projection-step kernels: operator + RHS build, matrix-free apply,
// pressure scatter/ghost fill, velocity correction.
//
// Coefficient and correction arithmetic keeps the reference's evaluation
// order with FP contraction off, so the assembled rows are bit-identical to
// a line-by-line FP64 restatement of ProjectionP_Mod.f90.
#include "sg_internal.hpp"
#include "sg_reduce.hpp"

namespace sg {

namespace {

// Ghosted (0:ni+1,0:nj+1) and cell (1:ni,1:nj) indexing with 0-based
// interior cell indices (ic,jc) <-> Fortran (ic+1,jc+1).
struct FLay {
    int ni, nj;
    __device__ size_t g(int gi, int gj) const { return (size_t)gi + (size_t)gj * (size_t)(ni + 2); }
    __device__ size_t c(int ic, int jc) const { return (size_t)ic + (size_t)jc * (size_t)ni; }
};

__device__ inline bool is_wet(double vof, double phi, double vofS) {
    // ProjectionP_Mod.f90:409-410
    return (vof >= 0.5 && vofS < EPSI) || (phi < 0.0 && vofS >= EPSI);
}

struct CoefIn {
    FLay L;
    const double *vof, *phi, *vofS, *DpU, *DpV;
    const double *aE, *aW, *aN, *aS;
    const double *dx, *dy, *udx, *vdy;
    double BetaP, BetaM;
};

// PoCoef(i,j,1..4) of ComputePossionMatrixCoefficient (ProjectionP_Mod.f90:
// 395-580) for one cell.  The dry/dry south face (:504-508) divides by the
// BetaD left by the previous cell of the loop, which is the dry (i,j-1)'s
// north face with the dry (i,j): BetaP.
__device__ inline void cell_pocoef(const CoefIn& in, int ic, int jc, double pc[4]) {
#pragma clang fp contract(off)
    const FLay& L = in.L;
    const int ni = L.ni, nj = L.nj;
    const double tol = TOLB;
    const double BetaP = in.BetaP, BetaM = in.BetaM;
    const size_t k = L.c(ic, jc);
    const double phic = in.phi[k];
    const bool wc = is_wet(in.vof[k], phic, in.vofS[k]);
    const double dxi = in.dx[ic], dyj = in.dy[jc];
    pc[0] = pc[1] = pc[2] = pc[3] = 0.0;
    // neighbour classification and |phi|
    auto nb = [&](int i2, int j2, bool& w, double& aphi) {
        const size_t kk = L.c(i2, j2);
        w = is_wet(in.vof[kk], in.phi[kk], in.vofS[kk]);
        aphi = fabs(in.phi[kk]);
    };
    // face Beta for a two-phase face seen from this cell
    auto beta_face = [&](bool wn, double aphin, bool& two) -> double {
        two = (wc != wn);
        if (!two) return wc ? BetaM : BetaP;
        const double Lamda = fabs(phic) / (fabs(phic) + aphin + tol);
        return wc ? (Lamda * BetaM + (1.0 - Lamda) * BetaP) : (Lamda * BetaP + (1.0 - Lamda) * BetaM);
    };
    bool wn, two;
    double aphin, B;
    // South
    if (jc > 0) {
        nb(ic, jc - 1, wn, aphin);
        B = beta_face(wn, aphin, two);
        pc[0] = in.DpV[L.g(ic + 1, jc)] * dxi / in.vdy[jc - 1] * BetaP * BetaM / B;
    } else {
        B = wc ? BetaM : BetaP;
        pc[0] = in.DpV[L.g(ic + 1, jc + 1)] * dxi / in.vdy[jc] * BetaP * BetaM / B;
    }
    // West
    if (ic > 0) {
        nb(ic - 1, jc, wn, aphin);
        B = beta_face(wn, aphin, two);
        pc[1] = in.DpU[L.g(ic, jc + 1)] * dyj / in.udx[ic - 1] * BetaP * BetaM / B;
    } else {
        B = wc ? BetaM : BetaP;
        pc[1] = in.DpU[L.g(ic + 1, jc + 1)] * dyj / in.udx[ic] * BetaP * BetaM / B;
    }
    // East
    if (ic < ni - 1) {
        nb(ic + 1, jc, wn, aphin);
        B = beta_face(wn, aphin, two);
    } else {
        B = wc ? BetaM : BetaP;
    }
    pc[2] = in.DpU[L.g(ic + 1, jc + 1)] * dyj / in.udx[ic] * BetaP * BetaM / B;
    // North
    if (jc < nj - 1) {
        nb(ic, jc + 1, wn, aphin);
        B = beta_face(wn, aphin, two);
    } else {
        B = wc ? BetaM : BetaP;
    }
    pc[3] = in.DpV[L.g(ic + 1, jc + 1)] * dxi / in.vdy[jc] * BetaP * BetaM / B;
}

struct RowOut {
    double S, W, D, E, N;
    bool fluid;
};

// One row of SetPoissonMatrix (ProjectionP_Mod.f90:167-266).
__device__ inline RowOut cell_row(const CoefIn& in, int ic, int jc, int bcW, int bcE, int bcS, int bcN) {
#pragma clang fp contract(off)
    const FLay& L = in.L;
    const int ni = L.ni, nj = L.nj;
    RowOut o{0.0, 0.0, 0.0, 0.0, 0.0, false};
    const double thr = 1.0 - EPSI;
    o.fluid = in.vofS[L.c(ic, jc)] < thr;
    if (!o.fluid) return o;
    double pc[4];
    cell_pocoef(in, ic, jc, pc);
    const size_t g = L.g(ic + 1, jc + 1);
    const double aS = in.aS[g], aW = in.aW[g], aE = in.aE[g], aN = in.aN[g];
    if (jc > 0 && in.vofS[L.c(ic, jc - 1)] < thr) o.S = -pc[0] * aS;
    if (ic > 0 && in.vofS[L.c(ic - 1, jc)] < thr) o.W = -pc[1] * aW;
    double d = pc[0] * aS + pc[1] * aW + pc[2] * aE + pc[3] * aN;
    if (bcS == 1 && jc == 0) d = d - pc[0] * aS;
    if (bcW == 1 && ic == 0) d = d - pc[1] * aW;
    if (bcE == 1 && ic == ni - 1) d = d - pc[2] * aE;
    if (bcN == 1 && jc == nj - 1) d = d - pc[3] * aN;
    d = d + (d < 0.0 ? -1.0 : 1.0) * TOLB;
    o.D = d;
    if (ic < ni - 1 && in.vofS[L.c(ic + 1, jc)] < thr) o.E = -pc[2] * aE;
    if (jc < nj - 1 && in.vofS[L.c(ic, jc + 1)] < thr) o.N = -pc[3] * aN;
    return o;
}

__device__ inline bool decoupled_row(const RowOut& o) {
    return o.fluid && o.S == 0.0 && o.W == 0.0 && o.E == 0.0 && o.N == 0.0 && fabs(o.D) <= TOLB;
}

struct RhsIn {
    FLay L;
    const double *u, *v, *aE, *aW, *aN, *aS, *nyS, *WlLh, *vofS, *dx, *dy;
    double vb;
};

// SetPoissonVectors (ProjectionP_Mod.f90:313-317)
__device__ inline double cell_rhs(const RhsIn& in, int ic, int jc) {
#pragma clang fp contract(off)
    const FLay& L = in.L;
    const size_t g = L.g(ic + 1, jc + 1);
    const size_t gw = L.g(ic, jc + 1), gs = L.g(ic + 1, jc);
    const size_t k = L.c(ic, jc);
    return -(in.u[g] * in.aE[g] - in.u[gw] * in.aW[g]) * in.dy[jc] -
           (in.v[g] * in.aN[g] - in.v[gs] * in.aS[g]) * in.dx[ic] + in.vb * in.nyS[k] * in.WlLh[k];
}

__global__ void __launch_bounds__(256) k_build(CoefIn cin, RhsIn rin, int bcW, int bcE, int bcS, int bcN, Dims D,
                                               double* __restrict__ diag, double* __restrict__ cE,
                                               double* __restrict__ cN, double* __restrict__ rhs,
                                               float* __restrict__ fdiag, float* __restrict__ fcE,
                                               float* __restrict__ fcN, int* guard) {
    const int ic = blockIdx.x * TX + threadIdx.x;
    const int jc = blockIdx.y * TY + threadIdx.y;
    if (ic >= D.ni || jc >= D.nj) return;
    const RowOut o = cell_row(cin, ic, jc, bcW, bcE, bcS, bcN);
    double b = 0.0;
    if (o.fluid) {
        b = cell_rhs(rin, ic, jc);
        int gbits = 0;
        if (isnan(o.D) || isnan(b)) gbits |= 1;
        if (fabs(o.D) > 1.0e10 || fabs(b) > 1.0e5) gbits |= 2;
        if (gbits) atomicOr(guard, gbits);
    }
    const size_t p = D.off(ic, jc);
    // a fluid cell whose four faces the cut-cell rules all closed keeps the
    // reference's row D = +1e-24 with no coupling (ProjectionP_Mod.f90:227):
    // it is decoupled, its value is b / D, set by the scatter; the solver
    // sees it as a solid row
    if (decoupled_row(o)) b = 0.0;
    // symmetric storage: the E/N entries of this row are the face couplings
    const double dd = (o.fluid && !decoupled_row(o)) ? o.D : 0.0;
    const double ce = -o.E, cn = -o.N;
    diag[p] = dd;
    cE[p] = ce;
    cN[p] = cn;
    rhs[p] = b;
    if (fdiag) {
        fdiag[p] = (float)dd;
        fcE[p] = (float)ce;
        fcN[p] = (float)cn;
    }
}

// right-hand side on the rows of the built operator (decoupled rows, whose
// diagonal is stored as 0, keep b = 0 as in k_build)
__global__ void __launch_bounds__(256) k_rhs(RhsIn rin, Dims D, const double* __restrict__ diag,
                                             double* __restrict__ rhs) {
    const int ic = blockIdx.x * TX + threadIdx.x;
    const int jc = blockIdx.y * TY + threadIdx.y;
    if (ic >= D.ni || jc >= D.nj) return;
    const bool fluid = rin.vofS[rin.L.c(ic, jc)] < 1.0 - EPSI;
    const size_t p = D.off(ic, jc);
    rhs[p] = (fluid && diag[p] != 0.0) ? cell_rhs(rin, ic, jc) : 0.0;
}

// SetPoissonVectors' b of every fluid row in the cell layout, decoupled
// rows included (the reference's vector; the solver's copy keeps 0 there)
__global__ void __launch_bounds__(256) k_rhs_cells(RhsIn rin, double* __restrict__ out) {
    const int ic = blockIdx.x * TX + threadIdx.x;
    const int jc = blockIdx.y * TY + threadIdx.y;
    if (ic >= rin.L.ni || jc >= rin.L.nj) return;
    const size_t k = rin.L.c(ic, jc);
    out[k] = rin.vofS[k] < 1.0 - EPSI ? cell_rhs(rin, ic, jc) : 0.0;
}

__global__ void __launch_bounds__(256) k_rows(CoefIn cin, int bcW, int bcE, int bcS, int bcN, double* __restrict__ matr) {
    const int ic = blockIdx.x * TX + threadIdx.x;
    const int jc = blockIdx.y * TY + threadIdx.y;
    const int ni = cin.L.ni, nj = cin.L.nj;
    if (ic >= ni || jc >= nj) return;
    const RowOut o = cell_row(cin, ic, jc, bcW, bcE, bcS, bcN);
    const size_t k = (size_t)ic + (size_t)jc * ni, N = (size_t)ni * nj;
    matr[k] = o.S;
    matr[k + N] = o.W;
    matr[k + 2 * N] = o.fluid ? o.D : 0.0;
    matr[k + 3 * N] = o.E;
    matr[k + 4 * N] = o.N;
}

__global__ void __launch_bounds__(256) k_apply5(Dims D, const double* __restrict__ diag, const double* __restrict__ cE,
                                                const double* __restrict__ cN, const double* __restrict__ x,
                                                double* __restrict__ y) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= D.ni || j >= D.nj) return;
    const size_t p = D.off(i, j);
    const size_t ld = D.ld;
    y[p] = diag[p] * x[p] - cE[p - 1] * x[p - 1] - cE[p] * x[p + 1] - cN[p - ld] * x[p - ld] - cN[p] * x[p + ld];
}

// DeltaPressureGetValues + DeltaPressureBoundaryCondition
// (ProjectionP_Mod.f90:377-391,582-595) on the ghosted Pp.
// A decoupled fluid row (all faces closed) takes the exact solution of its
// own equation, b / D.
__global__ void __launch_bounds__(256) k_scatter(Dims D, FLay L, const double* __restrict__ x,
                                                 const double* __restrict__ vofS, const double* __restrict__ diag,
                                                 CoefIn cin, RhsIn rin, int bcW, int bcE, int bcS, int bcN,
                                                 double* __restrict__ Pp) {
#pragma clang fp contract(off)
    const int gi = blockIdx.x * TX + threadIdx.x;
    const int gj = blockIdx.y * TY + threadIdx.y;
    const int ni = D.ni, nj = D.nj;
    if (gi > ni + 1 || gj > nj + 1) return;
    auto val = [&](int ic, int jc) -> double {
        if (!(vofS[L.c(ic, jc)] < 1.0 - EPSI)) return 0.0;
        const size_t p = D.off(ic, jc);
        if (diag[p] != 0.0) return x[p];
        const RowOut o = cell_row(cin, ic, jc, bcW, bcE, bcS, bcN);
        return decoupled_row(o) ? cell_rhs(rin, ic, jc) / o.D : x[p];
    };
    double out = 0.0;
    const bool in_i = gi >= 1 && gi <= ni, in_j = gj >= 1 && gj <= nj;
    if (in_i && in_j) out = val(gi - 1, gj - 1);
    else if (in_i && gj == 0) out = (double)bcS * val(gi - 1, 0);
    else if (in_i && gj == nj + 1) out = (double)bcN * val(gi - 1, nj - 1);
    else if (in_j && gi == 0) out = (double)bcW * val(0, gj - 1);
    else if (in_j && gi == ni + 1) out = (double)bcE * val(ni - 1, gj - 1);
    Pp[L.g(gi, gj)] = out;
}

struct CorrIn {
    FLay L;
    const double *vof, *phi, *vofS, *nyS, *WlLh, *aE, *aW, *aN, *aS;
    const int *uflag, *vflag;
    const double *vofSU, *vofSV, *DpU, *DpV, *upred, *vpred, *Pp;
    const double *dx, *dy, *udx, *vdy;
    double vb, BetaP, BetaM;
};

// ComputePUV_Mod.f90:59-100 (u) and :111-152 (v): Beta of the face from the
// wet/dry states of the two cells.
__device__ inline double corr_beta(const CorrIn& in, size_t k0, size_t k1) {
#pragma clang fp contract(off)
    const bool w0 = is_wet(in.vof[k0], in.phi[k0], in.vofS[k0]);
    const bool w1 = is_wet(in.vof[k1], in.phi[k1], in.vofS[k1]);
    const double BetaP = in.BetaP, BetaM = in.BetaM;
    if (w0 == w1) return w0 ? BetaM : BetaP;
    const double Lamda = fabs(in.phi[k0]) / (fabs(in.phi[k0]) + fabs(in.phi[k1]) + TOLB);
    return w0 ? (Lamda * BetaM + (1.0 - Lamda) * BetaP) : (Lamda * BetaP + (1.0 - Lamda) * BetaM);
}

__global__ void __launch_bounds__(256) k_correct_uv(CorrIn in, double* __restrict__ u, double* __restrict__ v,
                                                    double* __restrict__ p) {
#pragma clang fp contract(off)
    const int ic = blockIdx.x * TX + threadIdx.x;
    const int jc = blockIdx.y * TY + threadIdx.y;
    const FLay& L = in.L;
    const int ni = L.ni, nj = L.nj;
    if (ic >= ni || jc >= nj) return;
    const size_t k = L.c(ic, jc);
    const size_t g = L.g(ic + 1, jc + 1);
    const double BetaP = in.BetaP, BetaM = in.BetaM;
    if (ic < ni - 1 && in.uflag[k]) {
        const double B = corr_beta(in, k, L.c(ic + 1, jc));
        const double grad = (in.Pp[L.g(ic + 2, jc + 1)] - in.Pp[g]) * BetaM * BetaP / B * in.DpU[g] / in.udx[ic];
        u[g] = in.upred[g] - grad;
    }
    if (jc < nj - 1 && in.vflag[k]) {
        const double B = corr_beta(in, k, L.c(ic, jc + 1));
        const double grad = (in.Pp[L.g(ic + 1, jc + 2)] - in.Pp[g]) * BetaM * BetaP / B * in.DpV[g] / in.vdy[jc];
        v[g] = in.vpred[g] - grad;
    }
    if (in.vofS[k] < 1.0 - EPSI) p[g] = in.Pp[g];
}

// top-row v from continuity (ComputePUV_Mod.f90:177-182) and the mass
// residual (:184-217)
__global__ void __launch_bounds__(256) k_correct_mres(CorrIn in, const double* __restrict__ u, double* __restrict__ v,
                                                      double* __restrict__ mres) {
#pragma clang fp contract(off)
    const int ic = blockIdx.x * TX + threadIdx.x;
    const int jc = blockIdx.y * TY + threadIdx.y;
    const FLay& L = in.L;
    const int ni = L.ni, nj = L.nj;
    if (ic >= ni || jc >= nj) return;
    const size_t g = L.g(ic + 1, jc + 1), gw = L.g(ic, jc + 1), gs = L.g(ic + 1, jc);
    const double dxi = in.dx[ic], dyj = in.dy[jc];
    double vtop = v[g];
    if (jc == nj - 1) {
        vtop = ((-u[g] * in.aE[g] + u[gw] * in.aW[g]) * dyj + v[gs] * in.aS[g] * dxi) / dxi / in.aN[g];
        v[g] = vtop;
    }
    const size_t k = L.c(ic, jc);
    double m = (u[g] * in.aE[g] - u[gw] * in.aW[g]) * dyj + (vtop * in.aN[g] - v[gs] * in.aS[g]) * dxi -
               in.vb * in.nyS[k] * in.WlLh[k];
    if (in.vofS[k] > 1.0 - EPSI) m = 0.0;
    mres[k] = m;
}

// VariablesInternalCellCondition (ComputePUV_Mod.f90:227-245)
__global__ void __launch_bounds__(256) k_internal(CorrIn in, double* __restrict__ u, double* __restrict__ v,
                                                  double* __restrict__ p) {
    const int ic = blockIdx.x * TX + threadIdx.x;
    const int jc = blockIdx.y * TY + threadIdx.y;
    const FLay& L = in.L;
    if (ic >= L.ni || jc >= L.nj) return;
    const size_t k = L.c(ic, jc);
    const size_t g = L.g(ic + 1, jc + 1);
    if (in.vofS[k] > 1.0 - EPSI) p[g] = 0.0;
    if (in.vofSU[k] >= 1.0 - EPSI) u[g] = 0.0;
    if (in.vofSV[k] >= 1.0 - EPSI) v[g] = 0.0;
}

__global__ void __launch_bounds__(256) k_pad_to_cell(Dims D, const double* __restrict__ src, double* __restrict__ dst) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= D.ni || j >= D.nj) return;
    dst[(size_t)i + (size_t)j * D.ni] = src[D.off(i, j)];
}

__global__ void __launch_bounds__(256) k_cell_to_pad(Dims D, const double* __restrict__ src, double* __restrict__ dst) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= D.ni || j >= D.nj) return;
    dst[D.off(i, j)] = src[(size_t)i + (size_t)j * D.ni];
}

CoefIn make_coef_in(Ctx* c) {
    CoefIn in;
    in.L = FLay{c->ni, c->nj};
    in.vof = (const double*)c->field[SG_F_VOF];
    in.phi = (const double*)c->field[SG_F_PHI];
    in.vofS = (const double*)c->field[SG_F_VOFS];
    in.DpU = (const double*)c->field[SG_F_DPU];
    in.DpV = (const double*)c->field[SG_F_DPV];
    in.aE = (const double*)c->field[SG_F_AE];
    in.aW = (const double*)c->field[SG_F_AW];
    in.aN = (const double*)c->field[SG_F_AN];
    in.aS = (const double*)c->field[SG_F_AS];
    in.dx = c->dx;
    in.dy = c->dy;
    in.udx = c->udx;
    in.vdy = c->vdy;
    // ProjectionP_Mod.f90:405-406
    in.BetaP = 1.0 / (c->prm.rho_l / c->prm.rho_ref);
    in.BetaM = 1.0 / (c->prm.rho_g / c->prm.rho_ref);
    return in;
}

RhsIn make_rhs_in(Ctx* c) {
    RhsIn in;
    in.L = FLay{c->ni, c->nj};
    in.u = (const double*)c->field[SG_F_UPRED];
    in.v = (const double*)c->field[SG_F_VPRED];
    in.aE = (const double*)c->field[SG_F_AE];
    in.aW = (const double*)c->field[SG_F_AW];
    in.aN = (const double*)c->field[SG_F_AN];
    in.aS = (const double*)c->field[SG_F_AS];
    in.nyS = (const double*)c->field[SG_F_NYS];
    in.WlLh = (const double*)c->field[SG_F_WLLH];
    in.vofS = (const double*)c->field[SG_F_VOFS];
    in.dx = c->dx;
    in.dy = c->dy;
    in.vb = c->vb;
    return in;
}

}  // namespace

int poisson_build(Ctx* c) {
    Level& L0 = c->lev[0];
    const bool mixed = c->prm.precision == SG_PREC_MIXED;
    hipLaunchKernelGGL(k_build, grid2d(c->ni, c->nj), block2d(), 0, c->stream, make_coef_in(c), make_rhs_in(c),
                       c->prm.bc_w, c->prm.bc_e, c->prm.bc_s, c->prm.bc_n, L0.d, L0.A.diag, L0.A.cE, L0.A.cN,
                       c->rhs, mixed ? L0.Af.diag : nullptr, mixed ? L0.Af.cE : nullptr,
                       mixed ? L0.Af.cN : nullptr, &c->sc->guard);
    SG_HIP(hipGetLastError());
    c->built = 1;
    c->mg_ready = 0;
    return SG_OK;
}

int poisson_rhs_only(Ctx* c) {
    hipLaunchKernelGGL(k_rhs, grid2d(c->ni, c->nj), block2d(), 0, c->stream, make_rhs_in(c), c->lev[0].d,
                       (const double*)c->lev[0].A.diag, c->rhs);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int poisson_rhs_cells(Ctx* c, double* d_out) {
    hipLaunchKernelGGL(k_rhs_cells, grid2d(c->ni, c->nj), block2d(), 0, c->stream, make_rhs_in(c), d_out);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int poisson_matrix_rows(Ctx* c, double* d_matr) {
    hipLaunchKernelGGL(k_rows, grid2d(c->ni, c->nj), block2d(), 0, c->stream, make_coef_in(c), c->prm.bc_w,
                       c->prm.bc_e, c->prm.bc_s, c->prm.bc_n, d_matr);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int poisson_apply_fine(Ctx* c, const double* x, double* y) {
    Level& L0 = c->lev[0];
    hipLaunchKernelGGL(k_apply5, grid2d(c->ni, c->nj), block2d(), 0, c->stream, L0.d, L0.A.diag, L0.A.cE, L0.A.cN, x,
                       y);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int poisson_scatter(Ctx* c) {
    hipLaunchKernelGGL(k_scatter, grid2d(c->ni + 2, c->nj + 2), block2d(), 0, c->stream, c->lev[0].d,
                       FLay{c->ni, c->nj}, c->x, (const double*)c->field[SG_F_VOFS],
                       (const double*)c->lev[0].A.diag, make_coef_in(c), make_rhs_in(c), c->prm.bc_w,
                       c->prm.bc_e, c->prm.bc_s, c->prm.bc_n, (double*)c->field[SG_F_PP]);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int velocity_correct(Ctx* c) {
    CorrIn in;
    in.L = FLay{c->ni, c->nj};
    in.vof = (const double*)c->field[SG_F_VOF];
    in.phi = (const double*)c->field[SG_F_PHI];
    in.vofS = (const double*)c->field[SG_F_VOFS];
    in.nyS = (const double*)c->field[SG_F_NYS];
    in.WlLh = (const double*)c->field[SG_F_WLLH];
    in.aE = (const double*)c->field[SG_F_AE];
    in.aW = (const double*)c->field[SG_F_AW];
    in.aN = (const double*)c->field[SG_F_AN];
    in.aS = (const double*)c->field[SG_F_AS];
    in.uflag = (const int*)c->field[SG_F_UFLAG];
    in.vflag = (const int*)c->field[SG_F_VFLAG];
    in.vofSU = (const double*)c->field[SG_F_VOFSU];
    in.vofSV = (const double*)c->field[SG_F_VOFSV];
    in.DpU = (const double*)c->field[SG_F_DPU];
    in.DpV = (const double*)c->field[SG_F_DPV];
    in.upred = (const double*)c->field[SG_F_UPRED];
    in.vpred = (const double*)c->field[SG_F_VPRED];
    in.Pp = (const double*)c->field[SG_F_PP];
    in.dx = c->dx;
    in.dy = c->dy;
    in.udx = c->udx;
    in.vdy = c->vdy;
    in.vb = c->vb;
    in.BetaP = 1.0 / (c->prm.rho_l / c->prm.rho_ref);
    in.BetaM = 1.0 / (c->prm.rho_g / c->prm.rho_ref);
    double* u = (double*)c->field[SG_F_U];
    double* v = (double*)c->field[SG_F_V];
    double* p = (double*)c->field[SG_F_P];
    double* mres = (double*)c->field[SG_F_MRES];
    hipLaunchKernelGGL(k_correct_uv, grid2d(c->ni, c->nj), block2d(), 0, c->stream, in, u, v, p);
    SG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_correct_mres, grid2d(c->ni, c->nj), block2d(), 0, c->stream, in, u, v, mres);
    SG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_internal, grid2d(c->ni, c->nj), block2d(), 0, c->stream, in, u, v, p);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int padded_to_cell(Ctx* c, const double* src, double* dst_cell) {
    hipLaunchKernelGGL(k_pad_to_cell, grid2d(c->ni, c->nj), block2d(), 0, c->stream, c->lev[0].d, src, dst_cell);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int cell_to_padded(Ctx* c, const double* src_cell, double* dst) {
    hipLaunchKernelGGL(k_cell_to_pad, grid2d(c->ni, c->nj), block2d(), 0, c->stream, c->lev[0].d, src_cell, dst);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

}  // namespace sg
